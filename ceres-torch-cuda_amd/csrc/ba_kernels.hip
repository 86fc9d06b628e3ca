// Bundle-adjustment kernels for gfx950: per-observation evaluation, per-point
// elimination, camera-pair Schur assembly, step application and the
// per-observation candidate evaluation.  Templated on the assembly precision (float: mixed,
// double: plain FP64); parameters, residuals, costs and the model decrease are
// always FP64.
#include <algorithm>

#include "ba_kernels.h"
#include "ba_camera.h"
#include "lm_logic.h"

namespace ct {

// ---------------------------------------------------------------------------
// K1a: one thread per observation - residual, Jacobian, loss correction.
// Writes Jc [2x9], Jp [2x3] (assembly precision) and r (FP64); block partial
// of the cost.
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(kPointThreads, 2) void k_eval_obs(BAView v, BABuffers<real> b, BAWork w) {
  __shared__ double red[kPointThreads / kWave];
  const LMState* st = w.state;
  if (st->done) return;
  const int cur = st->cur;
  const double* __restrict__ cams = w.xcam + (size_t)cur * v.C * v.cam_stride;
  const double* __restrict__ pts = w.xpt + (size_t)cur * v.P * 3;
  const int o = blockIdx.x * kPointThreads + threadIdx.x;
  double cost = 0.0;
  if (o < v.N) {
    const int ci = v.obs_camera[o];
    const int p = v.obs_point[o];
    const double X[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    const double obs[2] = {v.obs_xy[2 * o], v.obs_xy[2 * o + 1]};
    double r[2], Jc[2][9], Jp[2][3];
    camera_jacobian(v, cams, ci, X, obs, r, Jc, Jp);
    if (v.camera_mask) {
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double m = v.camera_mask[9 * ci + j] ? 1.0 : 0.0;
        Jc[0][j] *= m;
        Jc[1][j] *= m;
      }
    }
    if (v.point_mask) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double m = v.point_mask[3 * p + j] ? 1.0 : 0.0;
        Jp[0][j] *= m;
        Jp[1][j] *= m;
      }
    }
    cost = robustify2(v.loss, r, Jc, Jp);
    b.r[2 * (size_t)o] = r[0];
    b.r[2 * (size_t)o + 1] = r[1];
    real* jc = b.JP + kJPStride * (size_t)o;
    real* jp = b.Jp + 6 * (size_t)o;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
#pragma unroll
      for (int j = 0; j < 9; ++j) jc[9 * k + j] = (real)Jc[k][j];
#pragma unroll
      for (int j = 0; j < 3; ++j) jp[3 * k + j] = (real)Jp[k][j];
    }
  }
  const double tc = block_sum<kPointThreads>(cost, red);
  if (threadIdx.x == 0) w.part_cost[blockIdx.x] = tc;
}

// ---------------------------------------------------------------------------
// K1b: one thread per point - point elimination.
//   V_p = sum Jp^T Jp + D_p, g_p = sum Jp^T r, L = chol(V_p),
//   q = L^-1 g_p, P_o = Jp_o L^-T, e_o = P_o q - r_o.
// Block partials: |g_p|^2, max|g_p|.
// ---------------------------------------------------------------------------
// A point without observations: its block of the system is the damping alone
// and its gradient is zero, so its step is zero.  L^-1 holds 1/sqrt(D), or 0
// where D is 0 (a free point under DOGLEG: the Gauss-Newton system carries no
// information on it, and its step stays zero).
template <typename real>
__device__ __forceinline__ void unobserved_point(const BAView& v, const BABuffers<real>& b, const BAWork& w,
                                                 const LMState* st, int p) {
  real inv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double d = column_damping(st, 0.0, !v.point_mask || v.point_mask[3 * p + k]);
    w.Dp[3 * p + k] = d;
    inv[k] = d > 0.0 ? real(1) / sqrt((real)d) : real(0);
    b.q[3 * p + k] = 0;
    b.gp[3 * p + k] = 0.0;
  }
  real* Li = b.Linv + 6 * (size_t)p;
  Li[0] = inv[0]; Li[1] = 0; Li[2] = inv[1]; Li[3] = 0; Li[4] = 0; Li[5] = inv[2];
}

template <typename real>
__global__ __launch_bounds__(kPointThreads, 2) void k_eval_points(BAView v, BABuffers<real> b, BAWork w) {
  __shared__ double red[kPointThreads / kWave];
  const LMState* st = w.state;
  if (st->done) return;
  const int p = blockIdx.x * kPointThreads + threadIdx.x;
  double g2 = 0.0, gmax = 0.0;
  if (p < v.P && v.point_offsets[p] == v.point_offsets[p + 1]) {
    unobserved_point(v, b, w, st, p);
  } else if (p < v.P) {
    const int o0 = v.point_offsets[p], o1 = v.point_offsets[p + 1];
    real V00 = 0, V10 = 0, V11 = 0, V20 = 0, V21 = 0, V22 = 0;
    double g0 = 0.0, g1 = 0.0, g2v = 0.0;
    for (int o = o0; o < o1; ++o) {
      const real* jp = b.Jp + 6 * (size_t)o;
      const real a0 = jp[0], a1 = jp[1], a2 = jp[2], c0 = jp[3], c1 = jp[4], c2 = jp[5];
      const double r0 = b.r[2 * (size_t)o], r1 = b.r[2 * (size_t)o + 1];
      V00 += a0 * a0 + c0 * c0;
      V10 += a1 * a0 + c1 * c0;
      V11 += a1 * a1 + c1 * c1;
      V20 += a2 * a0 + c2 * c0;
      V21 += a2 * a1 + c2 * c1;
      V22 += a2 * a2 + c2 * c2;
      g0 += (double)a0 * r0 + (double)c0 * r1;
      g1 += (double)a1 * r0 + (double)c1 * r1;
      g2v += (double)a2 * r0 + (double)c2 * r1;
    }
    // damping D = clamp(colsum J^2)/radius (linear.py:539-551)
    const double D0 = column_damping(st, (double)V00, !v.point_mask || v.point_mask[3 * p + 0]);
    const double D1 = column_damping(st, (double)V11, !v.point_mask || v.point_mask[3 * p + 1]);
    const double D2 = column_damping(st, (double)V22, !v.point_mask || v.point_mask[3 * p + 2]);
    w.Dp[3 * p] = D0;
    w.Dp[3 * p + 1] = D1;
    w.Dp[3 * p + 2] = D2;
    V00 += (real)D0;
    V11 += (real)D1;
    V22 += (real)D2;
    const real l00 = sqrt(V00);
    const real l10 = V10 / l00, l20 = V20 / l00;
    const real l11 = sqrt(V11 - l10 * l10);
    const real l21 = (V21 - l20 * l10) / l11;
    const real l22 = sqrt(V22 - l20 * l20 - l21 * l21);
    const real i00 = real(1) / l00, i11 = real(1) / l11, i22 = real(1) / l22;
    const real i10 = -l10 * i00 * i11;
    const real i21 = -l21 * i11 * i22;
    const real i20 = -(l20 * i00 + l21 * i10) * i22;
    const real q0 = i00 * (real)g0;
    const real q1 = i10 * (real)g0 + i11 * (real)g1;
    const real q2 = i20 * (real)g0 + i21 * (real)g1 + i22 * (real)g2v;
    real* Li = b.Linv + 6 * (size_t)p;
    Li[0] = i00; Li[1] = i10; Li[2] = i11; Li[3] = i20; Li[4] = i21; Li[5] = i22;
    b.q[3 * p] = q0;
    b.q[3 * p + 1] = q1;
    b.q[3 * p + 2] = q2;
    b.gp[3 * p] = g0;
    b.gp[3 * p + 1] = g1;
    b.gp[3 * p + 2] = g2v;
    for (int o = o0; o < o1; ++o) {
      const real* jp = b.Jp + 6 * (size_t)o;
      real* P = b.JP + kJPStride * (size_t)o + 18;
      double e[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const real j0 = jp[3 * k], j1 = jp[3 * k + 1], j2 = jp[3 * k + 2];
        // P = Jp L^-T : P[k][m] = sum_{l<=m} Jp[k][l] Linv[m][l]
        const real p0 = j0 * i00;
        const real p1 = j0 * i10 + j1 * i11;
        const real p2 = j0 * i20 + j1 * i21 + j2 * i22;
        P[3 * k] = p0;
        P[3 * k + 1] = p1;
        P[3 * k + 2] = p2;
        e[k] = (double)(p0 * q0 + p1 * q1 + p2 * q2) - b.r[2 * (size_t)o + k];
      }
      b.e[2 * (size_t)o] = (real)e[0];
      b.e[2 * (size_t)o + 1] = (real)e[1];
    }
    g2 = g0 * g0 + g1 * g1 + g2v * g2v;
    gmax = fmax(fabs(g0), fmax(fabs(g1), fabs(g2v)));
  }
  const double tg = block_sum<kPointThreads>(g2, red);
  const double tm = block_max<kPointThreads>(gmax, red);
  if (threadIdx.x == 0) {
    w.part_eval[2 * blockIdx.x] = tg;
    w.part_eval[2 * blockIdx.x + 1] = tm;
  }
}

// ---------------------------------------------------------------------------
// K1 (fused): one wavefront per group of whole points (at most kWave
// observations, built at plan creation), one lane per observation.
//   lane:      residual, Jacobian, loss correction (as k_eval_obs);
//   point:     the point's first lane sums its observations' J_p^T J_p and
//              J_p^T r terms from LDS in observation order and forms the
//              damping, the 3x3 factor, L^-1 and q (as k_eval_points);
//   lane:      P = Jp L^-T and e = P q - r, then the whole JP record
//              (Jc and P) is written at once.
// Same arithmetic as the two-kernel form; J_p and r stay in registers and
// no record is written in two parts.  Block partials: cost, |g_p|^2 and
// max|g_p| per workgroup.
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(kFusedThreads) void k_eval_fused(BAView v, BABuffers<real> b, BAWork w,
                                                              int num_groups, int unobserved_points) {
  constexpr int kWaves = kFusedThreads / kWave;
  __shared__ double red[kWaves];
  __shared__ double term[kWaves][9][kWave];   // per-observation V (6, assembly precision) and g (3) terms
  __shared__ double pdat[kWaves][9][kWave];   // at the point's first lane: L^-1 (6) and q (3)
  const LMState* st = w.state;
  if (st->done) return;
  const int cur = st->cur;
  const double* __restrict__ cams = w.xcam + (size_t)cur * v.C * v.cam_stride;
  const double* __restrict__ pts = w.xpt + (size_t)cur * v.P * 3;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int grp = blockIdx.x * kWaves + wave;
  int o = 0, p = 0, head = 0, count = 0;
  bool active = false;
  if (grp < num_groups) {
    const int o0 = v.point_offsets[w.point_groups[grp]], o1 = v.point_offsets[w.point_groups[grp + 1]];
    o = o0 + lane;
    active = o < o1;
  }
  double cost = 0.0, g2 = 0.0, gmax = 0.0;
  double r[2] = {0.0, 0.0}, Jc[2][9];
  real jq[2][3];
  if (active) {
    const int ci = v.obs_camera[o];
    p = v.obs_point[o];
    const int po0 = v.point_offsets[p];
    head = lane - (o - po0);
    count = v.point_offsets[p + 1] - po0;
    const double X[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    const double obs[2] = {v.obs_xy[2 * o], v.obs_xy[2 * o + 1]};
    double Jp[2][3];
    camera_jacobian(v, cams, ci, X, obs, r, Jc, Jp);
    if (v.camera_mask) {
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double m = v.camera_mask[9 * ci + j] ? 1.0 : 0.0;
        Jc[0][j] *= m;
        Jc[1][j] *= m;
      }
    }
    if (v.point_mask) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double m = v.point_mask[3 * p + j] ? 1.0 : 0.0;
        Jp[0][j] *= m;
        Jp[1][j] *= m;
      }
    }
    cost = robustify2(v.loss, r, Jc, Jp);
    b.r[2 * (size_t)o] = r[0];
    b.r[2 * (size_t)o + 1] = r[1];
    real* jp = b.Jp + 6 * (size_t)o;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        jq[k][j] = (real)Jp[k][j];
        jp[3 * k + j] = jq[k][j];
      }
    const real a0 = jq[0][0], a1 = jq[0][1], a2 = jq[0][2], c0 = jq[1][0], c1 = jq[1][1], c2 = jq[1][2];
    term[wave][0][lane] = (double)(real)(a0 * a0 + c0 * c0);
    term[wave][1][lane] = (double)(real)(a1 * a0 + c1 * c0);
    term[wave][2][lane] = (double)(real)(a1 * a1 + c1 * c1);
    term[wave][3][lane] = (double)(real)(a2 * a0 + c2 * c0);
    term[wave][4][lane] = (double)(real)(a2 * a1 + c2 * c1);
    term[wave][5][lane] = (double)(real)(a2 * a2 + c2 * c2);
    term[wave][6][lane] = (double)a0 * r[0] + (double)c0 * r[1];
    term[wave][7][lane] = (double)a1 * r[0] + (double)c1 * r[1];
    term[wave][8][lane] = (double)a2 * r[0] + (double)c2 * r[1];
  }
  __syncthreads();
  if (unobserved_points && grp < num_groups) {
    // the group's points without observations have no lane of their own
    const int q0 = w.point_groups[grp], q1 = w.point_groups[grp + 1];
    for (int q = q0 + lane; q < q1; q += kWave)
      if (v.point_offsets[q] == v.point_offsets[q + 1]) unobserved_point(v, b, w, st, q);
  }
  if (active && head == lane) {
    real V00 = 0, V10 = 0, V11 = 0, V20 = 0, V21 = 0, V22 = 0;
    double g0 = 0.0, g1 = 0.0, g2v = 0.0;
    for (int k = lane; k < lane + count; ++k) {
      V00 += (real)term[wave][0][k];
      V10 += (real)term[wave][1][k];
      V11 += (real)term[wave][2][k];
      V20 += (real)term[wave][3][k];
      V21 += (real)term[wave][4][k];
      V22 += (real)term[wave][5][k];
      g0 += term[wave][6][k];
      g1 += term[wave][7][k];
      g2v += term[wave][8][k];
    }
    // damping D = clamp(colsum J^2)/radius (linear.py:539-551)
    const double D0 = column_damping(st, (double)V00, !v.point_mask || v.point_mask[3 * p + 0]);
    const double D1 = column_damping(st, (double)V11, !v.point_mask || v.point_mask[3 * p + 1]);
    const double D2 = column_damping(st, (double)V22, !v.point_mask || v.point_mask[3 * p + 2]);
    w.Dp[3 * p] = D0;
    w.Dp[3 * p + 1] = D1;
    w.Dp[3 * p + 2] = D2;
    V00 += (real)D0;
    V11 += (real)D1;
    V22 += (real)D2;
    const real l00 = sqrt(V00);
    const real l10 = V10 / l00, l20 = V20 / l00;
    const real l11 = sqrt(V11 - l10 * l10);
    const real l21 = (V21 - l20 * l10) / l11;
    const real l22 = sqrt(V22 - l20 * l20 - l21 * l21);
    const real i00 = real(1) / l00, i11 = real(1) / l11, i22 = real(1) / l22;
    const real i10 = -l10 * i00 * i11;
    const real i21 = -l21 * i11 * i22;
    const real i20 = -(l20 * i00 + l21 * i10) * i22;
    const real q0 = i00 * (real)g0;
    const real q1 = i10 * (real)g0 + i11 * (real)g1;
    const real q2 = i20 * (real)g0 + i21 * (real)g1 + i22 * (real)g2v;
    real* Li = b.Linv + 6 * (size_t)p;
    Li[0] = i00; Li[1] = i10; Li[2] = i11; Li[3] = i20; Li[4] = i21; Li[5] = i22;
    b.q[3 * p] = q0;
    b.q[3 * p + 1] = q1;
    b.q[3 * p + 2] = q2;
    b.gp[3 * p] = g0;
    b.gp[3 * p + 1] = g1;
    b.gp[3 * p + 2] = g2v;
    pdat[wave][0][lane] = i00; pdat[wave][1][lane] = i10; pdat[wave][2][lane] = i11;
    pdat[wave][3][lane] = i20; pdat[wave][4][lane] = i21; pdat[wave][5][lane] = i22;
    pdat[wave][6][lane] = q0; pdat[wave][7][lane] = q1; pdat[wave][8][lane] = q2;
    g2 = g0 * g0 + g1 * g1 + g2v * g2v;
    gmax = fmax(fabs(g0), fmax(fabs(g1), fabs(g2v)));
  }
  __syncthreads();
  if (active) {
    const real i00 = (real)pdat[wave][0][head], i10 = (real)pdat[wave][1][head], i11 = (real)pdat[wave][2][head];
    const real i20 = (real)pdat[wave][3][head], i21 = (real)pdat[wave][4][head], i22 = (real)pdat[wave][5][head];
    const real q0 = (real)pdat[wave][6][head], q1 = (real)pdat[wave][7][head], q2 = (real)pdat[wave][8][head];
    real rec[kJPStride];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 9; ++j) rec[9 * k + j] = (real)Jc[k][j];
    double e[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const real j0 = jq[k][0], j1 = jq[k][1], j2 = jq[k][2];
      // P = Jp L^-T : P[k][m] = sum_{l<=m} Jp[k][l] Linv[m][l]
      const real p0 = j0 * i00;
      const real p1 = j0 * i10 + j1 * i11;
      const real p2 = j0 * i20 + j1 * i21 + j2 * i22;
      rec[18 + 3 * k] = p0;
      rec[18 + 3 * k + 1] = p1;
      rec[18 + 3 * k + 2] = p2;
      e[k] = (double)(p0 * q0 + p1 * q1 + p2 * q2) - r[k];
    }
    real* dst = b.JP + kJPStride * (size_t)o;
#pragma unroll
    for (int k = 0; k < kJPStride; ++k) dst[k] = rec[k];
    b.e[2 * (size_t)o] = (real)e[0];
    b.e[2 * (size_t)o + 1] = (real)e[1];
  }
  const double tc = block_sum<kFusedThreads>(cost, red);
  const double tg = block_sum<kFusedThreads>(g2, red);
  const double tm = block_max<kFusedThreads>(gmax, red);
  if (threadIdx.x == 0) {
    w.part_cost[blockIdx.x] = tc;
    w.part_eval[2 * blockIdx.x] = tg;
    w.part_eval[2 * blockIdx.x + 1] = tm;
  }
}

// ---------------------------------------------------------------------------
// K2a: one wavefront per work item (a bounded slice of one camera-pair block
// (i, j), i <= j):
//   kind 0 (Schur terms):  - sum Jc_a^T (P_a P_b^T) Jc_b
//   kind 1 (observations of camera i, diagonal blocks only):
//                          sum Jc^T Jc, and rhs = sum Jc^T e, g_c = sum Jc^T r,
//                          column norms diag(sum Jc^T Jc) in FP64.
// The slice is processed in rounds of kStageTerms terms.  Staging: lanes 2t and
// 2t+1 gather term t's two records (their own independent loads, all in flight
// together), swap P through a lane shuffle, and store Jc_a and
// T = (P_a P_b^T) Jc_b in LDS.  Accumulation: 63
// lanes form 7 groups of 3x3 sub-block owners; group g takes staged terms
// g, g+7, ...  The 7 partial blocks are summed in LDS in a fixed order and
// written as the item's partial.  Splitting blocks into items bounds the
// serial work per wavefront: a diagonal block covers every observation of
// its camera.
// ---------------------------------------------------------------------------
constexpr int kStageTerms = 32;  // terms per staging round (LDS per wavefront sets the occupancy)

template <typename real>
__global__ __launch_bounds__(kWave) void k_camera_items(BAView v, BABuffers<real> b, BAWork w) {
  // staging buffers during the rounds; the group partials reuse the space
  union Lds {
    struct {
      real L[kStageTerms][18];   // Jc_a of each staged term (kind 1: the observation's Jc)
      real R[kStageTerms][18];   // (P_a P_b^T) Jc_b (kind 1: Jc again)
      double V[kStageTerms][4];  // kind 1: e0, e1, r0, r1
    } stage;
    struct {
      real red[7][81];
      double vec[7][27];  // per group: rhs (9), g_c (9), column norms (9)
    } out;
  };
  __shared__ Lds lds;
  auto& sL = lds.stage.L;
  auto& sR = lds.stage.R;
  auto& sV = lds.stage.V;
  auto& red = lds.out.red;
  auto& vec = lds.out.vec;
  const LMState* st = w.state;
  if (st->done) return;
  const int slot = w.item_perm[blockIdx.x];
  const BlockItem it = w.items[slot];
  const int tid = threadIdx.x;
  const int grp = tid / 9;
  const int br = (tid % 9) / 3;
  const int bc = (tid % 9) % 3;
  const bool diag_obs = it.kind == 1;
  real acc[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[r][c] = 0;
  double rv[3] = {0.0, 0.0, 0.0}, gv[3] = {0.0, 0.0, 0.0};
  for (int base = it.begin; base < it.end; base += kStageTerms) {
    const int cnt = min(kStageTerms, it.end - base);
    if (diag_obs) {
      if (tid < cnt) {
        const int o = v.camera_obs[base + tid];
        const real* J = b.JP + kJPStride * (size_t)o;
#pragma unroll
        for (int k = 0; k < 18; ++k) {
          const real jv = J[k];
          sL[tid][k] = jv;
          sR[tid][k] = jv;
        }
        sV[tid][0] = (double)b.e[2 * (size_t)o];
        sV[tid][1] = (double)b.e[2 * (size_t)o + 1];
        sV[tid][2] = b.r[2 * (size_t)o];
        sV[tid][3] = b.r[2 * (size_t)o + 1];
      }
    } else {
      // lane pair (2t, 2t + 1) stages term t: the even lane gathers record a,
      // the odd lane record b, and the pair swaps P through a lane shuffle
      const int t = tid >> 1, side = tid & 1;
      const bool live = t < cnt;
      real jrec[18], pown[6];
      if (live) {
        const int o = v.block_terms[2 * (size_t)(base + t) + side];
        const real* rec = b.JP + kJPStride * (size_t)o;
#pragma unroll
        for (int k = 0; k < 18; ++k) jrec[k] = rec[k];
#pragma unroll
        for (int k = 0; k < 6; ++k) pown[k] = rec[18 + k];
      } else {
#pragma unroll
        for (int k = 0; k < 18; ++k) jrec[k] = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) pown[k] = 0;
      }
      real pother[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) pother[k] = __shfl_xor(pown[k], 1, kWave);
      if (live) {
        if (side == 0) {
#pragma unroll
          for (int k = 0; k < 18; ++k) sL[t][k] = jrec[k];
        } else {
          const real* pa = pother;
          const real* pb = pown;
          const real m00 = pa[0] * pb[0] + pa[1] * pb[1] + pa[2] * pb[2];
          const real m01 = pa[0] * pb[3] + pa[1] * pb[4] + pa[2] * pb[5];
          const real m10 = pa[3] * pb[0] + pa[4] * pb[1] + pa[5] * pb[2];
          const real m11 = pa[3] * pb[3] + pa[4] * pb[4] + pa[5] * pb[5];
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            const real b0 = jrec[k], b1 = jrec[9 + k];
            sR[t][k] = m00 * b0 + m01 * b1;
            sR[t][9 + k] = m10 * b0 + m11 * b1;
          }
        }
      }
    }
    __syncthreads();
    if (grp < 7) {
      for (int t = grp; t < cnt; t += 7) {
        real l0[3], l1[3], r0[3], r1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          l0[k] = sL[t][3 * br + k];
          l1[k] = sL[t][9 + 3 * br + k];
          r0[k] = sR[t][3 * bc + k];
          r1[k] = sR[t][9 + 3 * bc + k];
        }
        if (diag_obs) {
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[r][c] += l0[r] * r0[c] + l1[r] * r1[c];
          if (bc == 0) {
            const double e0 = sV[t][0], e1 = sV[t][1], q0 = sV[t][2], q1 = sV[t][3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              rv[r] += (double)l0[r] * e0 + (double)l1[r] * e1;
              gv[r] += (double)l0[r] * q0 + (double)l1[r] * q1;
            }
          }
        } else {
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[r][c] -= l0[r] * r0[c] + l1[r] * r1[c];
        }
      }
    }
    __syncthreads();
  }
  if (grp < 7) {
    if (diag_obs) {
      if (br == bc) {
#pragma unroll
        for (int r = 0; r < 3; ++r) vec[grp][18 + 3 * br + r] = (double)acc[r][r];
      }
      if (bc == 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          vec[grp][3 * br + r] = rv[r];
          vec[grp][9 + 3 * br + r] = gv[r];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) red[grp][(3 * br + r) * 9 + 3 * bc + c] = acc[r][c];
  }
  __syncthreads();
  real* out = b.item_S + 81 * (size_t)slot;
  for (int idx = tid; idx < 81; idx += kWave) {
    real s = 0;
    for (int gi = 0; gi < 7; ++gi) s += red[gi][idx];
    out[idx] = s;
  }
  if (diag_obs && tid < 27) {
    double s = 0.0;
    for (int gi = 0; gi < 7; ++gi) s += vec[gi][tid];
    w.item_vec[27 * (size_t)slot + tid] = s;
  }
}

// ---------------------------------------------------------------------------
// K2b: one thread per element of S, by 9x9 blocks in row-major order over the
// whole C x C block grid: sum the item partials of the block's camera pair
// (i, j), i <= j, in item order and write
//   S_ij = d_ij (U_i + D_i) - sum Jc_a^T (P_a P_b^T) Jc_b
// (a block below the diagonal takes the transposed element of its mirror),
// and, for diagonal blocks, rhs_i, g_c and D_i (linear.py:539-551).
// Consecutive workgroups fill consecutive stretches of the same rows of S,
// so every store line is written whole.  A diagonal block has one item per
// kItemMaxWork observations and per kItemMaxWork terms; the item loads are
// issued in batches ahead of their adds (which stay in item order).
// ---------------------------------------------------------------------------
constexpr int kReduceBlocks = 3;     // 9x9 blocks of S per workgroup, 81 threads each
constexpr int kReduceThreads = 256;  // the last 13 threads idle
static_assert(81 * kReduceBlocks <= kReduceThreads, "reduce workgroup too small");
constexpr int kSumBatch = 8;

// Ordered sums over a block's items [i0, i1), loads issued kSumBatch items
// ahead of the adds: s = sum item_S[81 k + e]; a = sum over the observation
// items (kind 1) of item_vec[27 k + ao] when ao >= 0.
template <typename real>
__device__ __forceinline__ void item_sums(const BABuffers<real>& b, const BAWork& w, int i0, int i1, int e, int ao,
                                          real& s, double& a) {
  s = 0;
  a = 0.0;
  for (int k0 = i0; k0 < i1; k0 += kSumBatch) {
    real v[kSumBatch];
    double va[kSumBatch];
    int kind[kSumBatch];
#pragma unroll
    for (int q = 0; q < kSumBatch; ++q) {
      const int k = k0 + q;
      const bool in = k < i1;
      v[q] = in ? b.item_S[81 * (size_t)k + e] : real(0);
      // every item has a vector slot; only the observation items' are summed
      kind[q] = in && ao >= 0 ? w.items[k].kind : 0;
      va[q] = in && ao >= 0 ? w.item_vec[27 * (size_t)k + ao] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kSumBatch; ++q) {
      if (k0 + q < i1) {
        s += v[q];
        if (kind[q] == 1) a += va[q];
      }
    }
  }
}

template <typename real>
__global__ __launch_bounds__(kReduceThreads) void k_camera_reduce(BAView v, BABuffers<real> b, BAWork w,
                                                                   int diagonal_only) {
  const LMState* st = w.state;
  if (st->done) return;
  const int C = v.C;
  // block of S: row-major over the C x C grid, or the C diagonal blocks only
  const long long ob = (long long)blockIdx.x * kReduceBlocks + threadIdx.x / 81;
  if (threadIdx.x >= 81 * kReduceBlocks || ob >= (diagonal_only ? (long long)C : (long long)C * C)) return;
  const int R = diagonal_only ? (int)ob : (int)(ob / C);
  const int Q = diagonal_only ? (int)ob : (int)(ob - (long long)R * C);
  const int e = threadIdx.x % 81;
  const int r = e / 9, c = e % 9;
  const int n = 9 * C;
  // camera pair (i, j), i <= j, and the element of its partials this thread sums
  const int i = R < Q ? R : Q, j = R < Q ? Q : R;
  const int blk = (int)((long long)i * C - (long long)i * (i - 1) / 2 + (j - i));  // row-major upper triangle
  const int i0 = w.block_item_offsets[blk], i1 = w.block_item_offsets[blk + 1];
  if (R != Q) {
    real s;
    double a;
    item_sums(b, w, i0, i1, R < Q ? e : 9 * c + r, -1, s, a);
    w.S[(size_t)(9 * R + r) * n + 9 * Q + c] = (double)s;
    return;
  }
  // symmetric: the lower triangle value.  Column r's damping comes from the
  // diagonal thread, its rhs and g_c from two off-diagonal threads of row r.
  const int role = r == c ? 0 : c == (r + 1) % 9 ? 1 : c == (r + 2) % 9 ? 2 : -1;
  const int ao = role == 0 ? 18 + r : role == 1 ? r : role == 2 ? 9 + r : -1;
  real s;
  double a;
  item_sums(b, w, i0, i1, r >= c ? e : 9 * c + r, ao, s, a);
  double val = (double)s;
  const int k = 9 * R + r;
  if (role == 0) {
    const double d = column_damping(st, a, !v.camera_mask || v.camera_mask[k]);
    w.Dc[k] = d;
    val += d;
  } else if (role == 1) {
    w.rhs[k] = a;
  } else if (role == 2) {
    w.gc[k] = a;
  }
  w.S[(size_t)k * n + 9 * R + c] = val;
}

// dx_p = -L^-T (q + sum_o P_o^T Jc_o dx_c)   (schur_solve_dense's xa, linear.py:698)
template <typename real>
__device__ __forceinline__ void backsub_point(const BAView& v, const BABuffers<real>& b, const double* __restrict__ dxc,
                                              int p, double* d0, double* d1, double* d2) {
  const int o0 = v.point_offsets[p], o1 = v.point_offsets[p + 1];
  const real* Li = b.Linv + 6 * (size_t)p;
  double s0 = (double)b.q[3 * p], s1 = (double)b.q[3 * p + 1], s2 = (double)b.q[3 * p + 2];
  for (int o = o0; o < o1; ++o) {
    const int ci = v.obs_camera[o];
    const real* J = b.JP + kJPStride * (size_t)o;
    double y0 = 0.0, y1 = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const double d = dxc[9 * ci + j];
      y0 += (double)J[j] * d;
      y1 += (double)J[9 + j] * d;
    }
    const real* P = b.JP + kJPStride * (size_t)o + 18;
    s0 += (double)P[0] * y0 + (double)P[3] * y1;
    s1 += (double)P[1] * y0 + (double)P[4] * y1;
    s2 += (double)P[2] * y0 + (double)P[5] * y1;
  }
  *d0 = -((double)Li[0] * s0 + (double)Li[1] * s1 + (double)Li[3] * s2);
  *d1 = -((double)Li[2] * s1 + (double)Li[4] * s2);
  *d2 = -((double)Li[5] * s2);
}

template <typename real>
__global__ __launch_bounds__(kPointThreads, 2) void k_backsub(BAView v, BABuffers<real> b, BAWork w) {
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int p = blockIdx.x * kPointThreads + threadIdx.x;
  if (p >= v.P) return;
  double d0, d1, d2;
  backsub_point(v, b, w.dxc, p, &d0, &d1, &d2);
  w.dxp[3 * p] = d0;
  w.dxp[3 * p + 1] = d1;
  w.dxp[3 * p + 2] = d2;
}

// ---------------------------------------------------------------------------
// Iterative refinement of (J^T J + D) dx = -g with the FP32 Jacobian and FP64
// accumulation (mixed precision).  One step:
//   rho = -g - (J^T J + D) dx                         (FP64)
//   delta_c = S^-1 (rho_c - sum_o Jc_o^T P_o L_p^-1 rho_p)
//   delta_p = L_p^-T (L_p^-1 rho_p - sum_o P_o^T Jc_o delta_c)
//   dx += delta
// reusing the point factors and the FP64 Cholesky factor of S.  The passes
// are fused across steps: the per-point pass of step k first completes the
// point's step (step 1: back-substitution from the solved dx_c; later steps:
// step k-1's update delta_p, with dx_c + delta_c standing for the updated
// camera step), then forms the point's part of rho; the per-camera pass
// applies the pending delta_c to its own camera before forming rho_c.  Only
// the last step's update is a pass of its own.  The formulas and their order
// are those of separate passes; the compiler's contraction into fused
// multiply-adds may differ in the last bit.
// ---------------------------------------------------------------------------
template <typename real, bool kFirst>
__global__ __launch_bounds__(kPointThreads, 2) void k_refine_pass(BAView v, BABuffers<real> b, BAWork w) {
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int p = blockIdx.x * kPointThreads + threadIdx.x;
  if (p >= v.P) return;
  const int o0 = v.point_offsets[p], o1 = v.point_offsets[p + 1];
  const real* Li = b.Linv + 6 * (size_t)p;
  double dp0, dp1, dp2;
  if constexpr (kFirst) {
    backsub_point(v, b, w.dxc, p, &dp0, &dp1, &dp2);
  } else {
    // the previous step's update of this point (its camera part waits for
    // the per-camera pass)
    double s0 = w.res_p[3 * p], s1 = w.res_p[3 * p + 1], s2 = w.res_p[3 * p + 2];
    for (int o = o0; o < o1; ++o) {
      const int ci = v.obs_camera[o];
      const real* J = b.JP + kJPStride * (size_t)o;
      double y0 = 0.0, y1 = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double d = w.dxc_corr[9 * ci + j];
        y0 += (double)J[j] * d;
        y1 += (double)J[9 + j] * d;
      }
      const real* P = b.JP + kJPStride * (size_t)o + 18;
      s0 -= (double)P[0] * y0 + (double)P[3] * y1;
      s1 -= (double)P[1] * y0 + (double)P[4] * y1;
      s2 -= (double)P[2] * y0 + (double)P[5] * y1;
    }
    dp0 = w.dxp[3 * p] + ((double)Li[0] * s0 + (double)Li[1] * s1 + (double)Li[3] * s2);
    dp1 = w.dxp[3 * p + 1] + ((double)Li[2] * s1 + (double)Li[4] * s2);
    dp2 = w.dxp[3 * p + 2] + (double)Li[5] * s2;
  }
  w.dxp[3 * p] = dp0;
  w.dxp[3 * p + 1] = dp1;
  w.dxp[3 * p + 2] = dp2;
  double rp0 = -b.gp[3 * p] - w.Dp[3 * p] * dp0;
  double rp1 = -b.gp[3 * p + 1] - w.Dp[3 * p + 1] * dp1;
  double rp2 = -b.gp[3 * p + 2] - w.Dp[3 * p + 2] * dp2;
  for (int o = o0; o < o1; ++o) {
    const int ci = v.obs_camera[o];
    const real* J = b.JP + kJPStride * (size_t)o;
    const real* Jq = b.Jp + 6 * (size_t)o;
    double u0 = (double)Jq[0] * dp0 + (double)Jq[1] * dp1 + (double)Jq[2] * dp2;
    double u1 = (double)Jq[3] * dp0 + (double)Jq[4] * dp1 + (double)Jq[5] * dp2;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const double d = kFirst ? w.dxc[9 * ci + j] : w.dxc[9 * ci + j] + w.dxc_corr[9 * ci + j];
      u0 += (double)J[j] * d;
      u1 += (double)J[9 + j] * d;
    }
    w.u_obs[4 * (size_t)o] = u0;
    w.u_obs[4 * (size_t)o + 1] = u1;
    rp0 -= (double)Jq[0] * u0 + (double)Jq[3] * u1;
    rp1 -= (double)Jq[1] * u0 + (double)Jq[4] * u1;
    rp2 -= (double)Jq[2] * u0 + (double)Jq[5] * u1;
  }
  const double t0 = (double)Li[0] * rp0;
  const double t1 = (double)Li[1] * rp0 + (double)Li[2] * rp1;
  const double t2 = (double)Li[3] * rp0 + (double)Li[4] * rp1 + (double)Li[5] * rp2;
  w.res_p[3 * p] = t0;
  w.res_p[3 * p + 1] = t1;
  w.res_p[3 * p + 2] = t2;
  for (int o = o0; o < o1; ++o) {
    const real* P = b.JP + kJPStride * (size_t)o + 18;
    w.u_obs[4 * (size_t)o + 2] = (double)P[0] * t0 + (double)P[1] * t1 + (double)P[2] * t2;
    w.u_obs[4 * (size_t)o + 3] = (double)P[3] * t0 + (double)P[4] * t1 + (double)P[5] * t2;
  }
}

// One wavefront per camera: rho_c and the reduced right-hand side (kUpdate:
// first the previous step's correction of this camera, dx_c += delta_c).
template <typename real, bool kUpdate>
__global__ __launch_bounds__(kWave) void k_refine_cameras(BAView v, BABuffers<real> b, BAWork w) {
  __shared__ double part[kWave][9];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int ci = blockIdx.x;
  const int tid = threadIdx.x;
  double acc[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) acc[j] = 0.0;
  for (int t = v.camera_offsets[ci] + tid; t < v.camera_offsets[ci + 1]; t += kWave) {
    const int o = v.camera_obs[t];
    const real* J = b.JP + kJPStride * (size_t)o;
    const double a0 = w.u_obs[4 * (size_t)o] + w.u_obs[4 * (size_t)o + 2];
    const double a1 = w.u_obs[4 * (size_t)o + 1] + w.u_obs[4 * (size_t)o + 3];
#pragma unroll
    for (int j = 0; j < 9; ++j) acc[j] += (double)J[j] * a0 + (double)J[9 + j] * a1;
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) part[tid][j] = acc[j];
  __syncthreads();
  if (tid < 9) {
    double ssum = 0.0;
    for (int l = 0; l < kWave; ++l) ssum += part[l][tid];
    const int k = 9 * ci + tid;
    double d = w.dxc[k];
    if (kUpdate) {
      d += w.dxc_corr[k];
      w.dxc[k] = d;
    }
    w.res_c[k] = -w.gc[k] - w.Dc[k] * d - ssum;
  }
}

template <typename real>
__global__ __launch_bounds__(kPointThreads, 2) void k_refine_update(BAView v, BABuffers<real> b, BAWork w) {
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int t = blockIdx.x * kPointThreads + threadIdx.x;
  if (t < v.P) {
    const int p = t;
    double s0 = w.res_p[3 * p], s1 = w.res_p[3 * p + 1], s2 = w.res_p[3 * p + 2];
    for (int o = v.point_offsets[p]; o < v.point_offsets[p + 1]; ++o) {
      const int ci = v.obs_camera[o];
      const real* J = b.JP + kJPStride * (size_t)o;
      double y0 = 0.0, y1 = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double d = w.dxc_corr[9 * ci + j];
        y0 += (double)J[j] * d;
        y1 += (double)J[9 + j] * d;
      }
      const real* P = b.JP + kJPStride * (size_t)o + 18;
      s0 -= (double)P[0] * y0 + (double)P[3] * y1;
      s1 -= (double)P[1] * y0 + (double)P[4] * y1;
      s2 -= (double)P[2] * y0 + (double)P[5] * y1;
    }
    const real* Li = b.Linv + 6 * (size_t)p;
    w.dxp[3 * p] += (double)Li[0] * s0 + (double)Li[1] * s1 + (double)Li[3] * s2;
    w.dxp[3 * p + 1] += (double)Li[2] * s1 + (double)Li[4] * s2;
    w.dxp[3 * p + 2] += (double)Li[5] * s2;
  }
  if (t < 9 * v.C) w.dxc[t] += w.dxc_corr[t];
}

// ---------------------------------------------------------------------------
// K5a: apply the step, x' = Plus(x, dx) (Euclidean Plus, manifolds.py:69-91;
// QuaternionManifold.plus on quaternion camera rotations), with
// the point back-substitution fused when the solve did not already do it.
// Block partials: |dx|^2, |x'|^2 over all parameters, non-finite count.
// ---------------------------------------------------------------------------
template <typename real, bool kBacksub>
__global__ __launch_bounds__(kPointThreads, 2) void k_step_points(BAView v, BABuffers<real> b, BAWork w) {
  __shared__ double red[kPointThreads / kWave];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int cur = st->cur;
  const double* __restrict__ cams = w.xcam + (size_t)cur * v.C * v.cam_stride;
  const double* __restrict__ pts = w.xpt + (size_t)cur * v.P * 3;
  double* __restrict__ ncams = w.xcam + (size_t)(1 - cur) * v.C * v.cam_stride;
  double* __restrict__ npts = w.xpt + (size_t)(1 - cur) * v.P * 3;
  const int t = blockIdx.x * kPointThreads + threadIdx.x;
  double dx2 = 0.0, x2 = 0.0, bad = 0.0;
  if (v.cam_model == CT_CAMERA_ANGLE_AXIS) {
    if (t < 9 * v.C) {
      const double d = w.dxc[t];
      const double nx = cams[t] + d;
      ncams[t] = nx;
      dx2 += d * d;
      x2 += nx * nx;
      if (!isfinite(d)) bad = 1.0;
    }
  } else if (t < v.C) {
    step_quat_camera(v, cams, ncams, w.dxc, t, dx2, x2, bad);
  }
  if (t < v.P) {
    double d0, d1, d2;
    if (kBacksub) {
      backsub_point(v, b, w.dxc, t, &d0, &d1, &d2);
      w.dxp[3 * t] = d0;
      w.dxp[3 * t + 1] = d1;
      w.dxp[3 * t + 2] = d2;
    } else {
      d0 = w.dxp[3 * t];
      d1 = w.dxp[3 * t + 1];
      d2 = w.dxp[3 * t + 2];
    }
    const double n0 = pts[3 * t] + d0, n1 = pts[3 * t + 1] + d1, n2 = pts[3 * t + 2] + d2;
    npts[3 * t] = n0;
    npts[3 * t + 1] = n1;
    npts[3 * t + 2] = n2;
    dx2 += d0 * d0 + d1 * d1 + d2 * d2;
    x2 += n0 * n0 + n1 * n1 + n2 * n2;
    if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) bad = 1.0;
  }
  const double a0 = block_sum<kPointThreads>(dx2, red);
  const double a1 = block_sum<kPointThreads>(x2, red);
  const double a2 = block_sum<kPointThreads>(bad, red);
  if (threadIdx.x == 0) {
    w.part_step[3 * blockIdx.x] = a0;
    w.part_step[3 * blockIdx.x + 1] = a1;
    w.part_step[3 * blockIdx.x + 2] = a2;
  }
}

// ---------------------------------------------------------------------------
// K5b: one thread per observation - candidate cost at x' and the model
// decrease term -r.u - u.u/2 with u = Jc dx_c + Jp dx_p (solver.py:798-802,
// summed per observation in FP64).
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(kPointThreads, 2) void k_candidate_obs(BAView v, BABuffers<real> b, BAWork w) {
  __shared__ double red[kPointThreads / kWave];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int cur = st->cur;
  const double* __restrict__ ncams = w.xcam + (size_t)(1 - cur) * v.C * v.cam_stride;
  const double* __restrict__ npts = w.xpt + (size_t)(1 - cur) * v.P * 3;
  const int o = blockIdx.x * kPointThreads + threadIdx.x;
  double cost = 0.0, md = 0.0;
  if (o < v.N) {
    const int ci = v.obs_camera[o];
    const int p = v.obs_point[o];
    const real* J = b.JP + kJPStride * (size_t)o;
    const real* Jq = b.Jp + 6 * (size_t)o;
    const double d0 = w.dxp[3 * p], d1 = w.dxp[3 * p + 1], d2 = w.dxp[3 * p + 2];
    double u0 = (double)Jq[0] * d0 + (double)Jq[1] * d1 + (double)Jq[2] * d2;
    double u1 = (double)Jq[3] * d0 + (double)Jq[4] * d1 + (double)Jq[5] * d2;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const double d = w.dxc[9 * ci + j];
      u0 += (double)J[j] * d;
      u1 += (double)J[9 + j] * d;
    }
    const double r0 = b.r[2 * (size_t)o], r1 = b.r[2 * (size_t)o + 1];
    md = -(r0 * u0 + r1 * u1) - 0.5 * (u0 * u0 + u1 * u1);
    const double X[3] = {npts[3 * p], npts[3 * p + 1], npts[3 * p + 2]};
    const double obs[2] = {v.obs_xy[2 * o], v.obs_xy[2 * o + 1]};
    double rr[2];
    camera_residual(v, ncams, ci, X, obs, rr);
    cost = robust_cost2(v.loss, rr);
  }
  const double a0 = block_sum<kPointThreads>(cost, red);
  const double a1 = block_sum<kPointThreads>(md, red);
  if (threadIdx.x == 0) {
    w.part_cand[2 * blockIdx.x] = a0;
    w.part_cand[2 * blockIdx.x + 1] = a1;
  }
}

// ---------------------------------------------------------------------------
// K5 (fused): step and candidate over the point groups of the fused
// evaluation (one wavefront per group, one lane per observation).
//   lane:  y = Jc dx_c and its back-substitution term P^T y;
//   point: the first lane forms dx_p = -L^-T (q + sum P^T y) in observation
//          order (or reads it when the refinement computed it), x'_p, and
//          its |dx|^2, |x'|^2 terms (as k_step_points);
//   lane:  u = Jp dx_p + Jc dx_c, the model-decrease term -r.u - u.u/2 and
//          the candidate cost at x' (as k_candidate_obs).
// The first 9C threads of the grid also step the cameras.  Each JP record is
// read once.  Block partials: |dx|^2, |x'|^2, non-finite count; candidate
// cost and model decrease, per workgroup.
// ---------------------------------------------------------------------------
template <typename real, bool kBacksub>
__global__ __launch_bounds__(kFusedThreads) void k_step_candidate(BAView v, BABuffers<real> b, BAWork w,
                                                                  int num_groups) {
  constexpr int kWaves = kFusedThreads / kWave;
  __shared__ double red[kWaves];
  __shared__ double term[kWaves][3][kWave];  // per observation: P^T (Jc dx_c)
  __shared__ double pnew[kWaves][6][kWave];  // at the point's first lane: dx_p (3), x'_p (3)
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int cur = st->cur;
  const double* __restrict__ cams = w.xcam + (size_t)cur * v.C * v.cam_stride;
  const double* __restrict__ pts = w.xpt + (size_t)cur * v.P * 3;
  double* __restrict__ ncams = w.xcam + (size_t)(1 - cur) * v.C * v.cam_stride;
  double* __restrict__ npts = w.xpt + (size_t)(1 - cur) * v.P * 3;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int grp = blockIdx.x * kWaves + wave;
  double dx2 = 0.0, x2 = 0.0, bad = 0.0, cost = 0.0, md = 0.0;
  // cameras: x'_c = Plus(x_c, dx_c)
  if (v.cam_model == CT_CAMERA_ANGLE_AXIS) {
    for (int t = blockIdx.x * kFusedThreads + threadIdx.x; t < 9 * v.C; t += gridDim.x * kFusedThreads) {
      const double dd = w.dxc[t];
      const double nx = cams[t] + dd;
      ncams[t] = nx;
      dx2 += dd * dd;
      x2 += nx * nx;
      if (!isfinite(dd)) bad = 1.0;
    }
  } else {
    for (int t = blockIdx.x * kFusedThreads + threadIdx.x; t < v.C; t += gridDim.x * kFusedThreads)
      step_quat_camera(v, cams, ncams, w.dxc, t, dx2, x2, bad);
  }
  int o = 0, p = 0, head = 0, count = 0, ci = 0;
  bool active = false;
  if (grp < num_groups) {
    const int o0 = v.point_offsets[w.point_groups[grp]], o1 = v.point_offsets[w.point_groups[grp + 1]];
    o = o0 + lane;
    active = o < o1;
  }
  real J[18];
  if (active) {
    ci = v.obs_camera[o];
    p = v.obs_point[o];
    const int po0 = v.point_offsets[p];
    head = lane - (o - po0);
    count = v.point_offsets[p + 1] - po0;
    const real* rec = b.JP + kJPStride * (size_t)o;
#pragma unroll
    for (int k = 0; k < 18; ++k) J[k] = rec[k];
    if (kBacksub) {
      double y0 = 0.0, y1 = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double dd = w.dxc[9 * ci + j];
        y0 += (double)J[j] * dd;
        y1 += (double)J[9 + j] * dd;
      }
      const real* P = rec + 18;
      term[wave][0][lane] = (double)P[0] * y0 + (double)P[3] * y1;
      term[wave][1][lane] = (double)P[1] * y0 + (double)P[4] * y1;
      term[wave][2][lane] = (double)P[2] * y0 + (double)P[5] * y1;
    }
  }
  __syncthreads();
  if (active && head == lane) {
    double d0, d1, d2;
    if (kBacksub) {
      // dx_p = -L^-T (q + sum_o P_o^T Jc_o dx_c)   (schur_solve_dense's xa, linear.py:698)
      const real* Li = b.Linv + 6 * (size_t)p;
      double s0 = (double)b.q[3 * p], s1 = (double)b.q[3 * p + 1], s2 = (double)b.q[3 * p + 2];
      for (int k = lane; k < lane + count; ++k) {
        s0 += term[wave][0][k];
        s1 += term[wave][1][k];
        s2 += term[wave][2][k];
      }
      d0 = -((double)Li[0] * s0 + (double)Li[1] * s1 + (double)Li[3] * s2);
      d1 = -((double)Li[2] * s1 + (double)Li[4] * s2);
      d2 = -((double)Li[5] * s2);
      w.dxp[3 * p] = d0;
      w.dxp[3 * p + 1] = d1;
      w.dxp[3 * p + 2] = d2;
    } else {
      d0 = w.dxp[3 * p];
      d1 = w.dxp[3 * p + 1];
      d2 = w.dxp[3 * p + 2];
    }
    const double n0 = pts[3 * p] + d0, n1 = pts[3 * p + 1] + d1, n2 = pts[3 * p + 2] + d2;
    npts[3 * p] = n0;
    npts[3 * p + 1] = n1;
    npts[3 * p + 2] = n2;
    dx2 += d0 * d0 + d1 * d1 + d2 * d2;
    x2 += n0 * n0 + n1 * n1 + n2 * n2;
    if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) bad = 1.0;
    pnew[wave][0][lane] = d0; pnew[wave][1][lane] = d1; pnew[wave][2][lane] = d2;
    pnew[wave][3][lane] = n0; pnew[wave][4][lane] = n1; pnew[wave][5][lane] = n2;
  }
  __syncthreads();
  if (active) {
    const real* Jq = b.Jp + 6 * (size_t)o;
    const double d0 = pnew[wave][0][head], d1 = pnew[wave][1][head], d2 = pnew[wave][2][head];
    double u0 = (double)Jq[0] * d0 + (double)Jq[1] * d1 + (double)Jq[2] * d2;
    double u1 = (double)Jq[3] * d0 + (double)Jq[4] * d1 + (double)Jq[5] * d2;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const double dd = w.dxc[9 * ci + j];
      u0 += (double)J[j] * dd;
      u1 += (double)J[9 + j] * dd;
    }
    const double r0 = b.r[2 * (size_t)o], r1 = b.r[2 * (size_t)o + 1];
    md = -(r0 * u0 + r1 * u1) - 0.5 * (u0 * u0 + u1 * u1);
    const double X[3] = {pnew[wave][3][head], pnew[wave][4][head], pnew[wave][5][head]};
    const double obs[2] = {v.obs_xy[2 * o], v.obs_xy[2 * o + 1]};
    double rr[2];
    camera_residual_after_step(v, cams, ci, w.dxc + 9 * (size_t)ci, X, obs, rr);
    cost = robust_cost2(v.loss, rr);
  }
  const double a0 = block_sum<kFusedThreads>(dx2, red);
  const double a1 = block_sum<kFusedThreads>(x2, red);
  const double a2 = block_sum<kFusedThreads>(bad, red);
  const double a3 = block_sum<kFusedThreads>(cost, red);
  const double a4 = block_sum<kFusedThreads>(md, red);
  if (threadIdx.x == 0) {
    w.part_step[3 * blockIdx.x] = a0;
    w.part_step[3 * blockIdx.x + 1] = a1;
    w.part_step[3 * blockIdx.x + 2] = a2;
    w.part_cand[2 * blockIdx.x] = a3;
    w.part_cand[2 * blockIdx.x + 1] = a4;
  }
}

// ---------------------------------------------------------------------------
// DOGLEG step (linear.py:554-608).  With the Gauss-Newton step gn = (dxc, dxp)
// of the undamped Schur system and the gradient g = (gc, gp):
//   k_dogleg_partials: per workgroup |g|^2, |gn|^2, g.gn (cameras: the first 9C
//     threads of the grid; points: one thread each) and, over each point's
//     observations, |J g|^2, |J gn|^2, (J g).(J gn);
//   k_dogleg_coefficients: the fixed-order sum and the step coefficients;
//   k_dogleg_combine: dx <- cgn gn + cg g, in place.
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(kPointThreads, 2) void k_dogleg_partials(BAView v, BABuffers<real> b, BAWork w) {
  __shared__ double red[kPointThreads / kWave];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  double d[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = blockIdx.x * kPointThreads + threadIdx.x; t < 9 * v.C; t += gridDim.x * kPointThreads) {
    const double gg = w.gc[t], nn = w.dxc[t];
    d[0] += gg * gg;
    d[1] += nn * nn;
    d[2] += gg * nn;
  }
  const int p = blockIdx.x * kPointThreads + threadIdx.x;
  if (p < v.P) {
    const double g0 = b.gp[3 * p], g1 = b.gp[3 * p + 1], g2 = b.gp[3 * p + 2];
    const double n0 = w.dxp[3 * p], n1 = w.dxp[3 * p + 1], n2 = w.dxp[3 * p + 2];
    d[0] += g0 * g0 + g1 * g1 + g2 * g2;
    d[1] += n0 * n0 + n1 * n1 + n2 * n2;
    d[2] += g0 * n0 + g1 * n1 + g2 * n2;
    const int o0 = v.point_offsets[p], o1 = v.point_offsets[p + 1];
    for (int o = o0; o < o1; ++o) {
      const int ci = v.obs_camera[o];
      const real* J = b.JP + kJPStride * (size_t)o;
      const real* Jq = b.Jp + 6 * (size_t)o;
      double jg0 = (double)Jq[0] * g0 + (double)Jq[1] * g1 + (double)Jq[2] * g2;
      double jg1 = (double)Jq[3] * g0 + (double)Jq[4] * g1 + (double)Jq[5] * g2;
      double jn0 = (double)Jq[0] * n0 + (double)Jq[1] * n1 + (double)Jq[2] * n2;
      double jn1 = (double)Jq[3] * n0 + (double)Jq[4] * n1 + (double)Jq[5] * n2;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double gc = w.gc[9 * ci + j], nc = w.dxc[9 * ci + j];
        jg0 += (double)J[j] * gc;
        jg1 += (double)J[9 + j] * gc;
        jn0 += (double)J[j] * nc;
        jn1 += (double)J[9 + j] * nc;
      }
      d[3] += jg0 * jg0 + jg1 * jg1;
      d[4] += jn0 * jn0 + jn1 * jn1;
      d[5] += jg0 * jn0 + jg1 * jn1;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double t = block_sum<kPointThreads>(d[k], red);
    if (threadIdx.x == 0) w.part_dogleg[6 * (size_t)blockIdx.x + k] = t;
  }
}

constexpr int kDoglegThreads = 256;
__global__ __launch_bounds__(kDoglegThreads) void k_dogleg_coefficients(BAWork w, int num_blocks) {
  __shared__ double red[kDoglegThreads / kWave];
  LMState* st = w.state;
  if (st->done || st->skip_step) return;
  double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < num_blocks; i += kDoglegThreads)
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] += w.part_dogleg[6 * (size_t)i + k];
  double d[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) d[k] = block_sum<kDoglegThreads>(a[k], red);
  if (threadIdx.x == 0) lm_dogleg_coefficients(st, d);
}

template <typename real>
__global__ __launch_bounds__(kPointThreads) void k_dogleg_combine(BAView v, BABuffers<real> b, BAWork w) {
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const double cgn = st->dl_cgn, cg = st->dl_cg;
  const int stride = gridDim.x * kPointThreads;
  for (int t = blockIdx.x * kPointThreads + threadIdx.x; t < 9 * v.C; t += stride) w.dxc[t] = cgn * w.dxc[t] + cg * w.gc[t];
  for (int t = blockIdx.x * kPointThreads + threadIdx.x; t < 3 * v.P; t += stride) w.dxp[t] = cgn * w.dxp[t] + cg * b.gp[t];
}

// Test hook (ct_lm_options.debug_invalid_steps): overwrite this iteration's
// whole step with NaN, as a linear solver returning NaN would leave it; the
// step kernels' finiteness test then rejects it.
__global__ __launch_bounds__(kPointThreads) void k_poison_step(BAWork w, int n_cam, int n_pt) {
  const LMState* st = w.state;
  if (st->done || st->skip_step || !st->poison) return;
  const double nan = __builtin_nan("");
  const int stride = gridDim.x * kPointThreads;
  for (int t = blockIdx.x * kPointThreads + threadIdx.x; t < n_cam; t += stride) w.dxc[t] = nan;
  for (int t = blockIdx.x * kPointThreads + threadIdx.x; t < n_pt; t += stride) w.dxp[t] = nan;
}

// Test hook: fill this workgroup's whole LDS allocation with v (volatile
// stores, so nothing elides them).  Launched with the largest LDS a workgroup
// may take, one workgroup holds a compute unit's LDS; two per compute unit
// cover every one of them.
__global__ __launch_bounds__(256) void k_fill_lds(double v, int count) {
  extern __shared__ double lds_fill[];
  volatile double* b = lds_fill;
  for (int i = threadIdx.x; i < count; i += blockDim.x) b[i] = v;
  __syncthreads();
}

// Plain evaluation (cost + corrected residuals) at given parameters.
__global__ __launch_bounds__(kPointThreads) void k_evaluate_only(BAView v, const double* cams, const double* pts,
                                                                  double* residuals, double* part) {
  __shared__ double red[kPointThreads / kWave];
  const int p = blockIdx.x * kPointThreads + threadIdx.x;
  double cost = 0.0;
  if (p < v.P) {
    const double X[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    for (int o = v.point_offsets[p]; o < v.point_offsets[p + 1]; ++o) {
      const int ci = v.obs_camera[o];
      const double obs[2] = {v.obs_xy[2 * o], v.obs_xy[2 * o + 1]};
      double r[2];
      camera_residual(v, cams, ci, X, obs, r);
      const double s = r[0] * r[0] + r[1] * r[1];
      double rho1, rho2;
      const double rho0 = loss_rho(v.loss, s, &rho1, &rho2);
      cost += 0.5 * rho0;
      if (residuals) {
        double sq = sqrt(fmax(rho1, 0.0)), scale = sq;
        if (s > 0.0 && rho2 > 0.0) scale = sq / sqrt(1.0 + 2.0 * s * rho2 / rho1);
        residuals[2 * o] = scale * r[0];
        residuals[2 * o + 1] = scale * r[1];
      }
    }
  }
  const double tc = block_sum<kPointThreads>(cost, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tc;
}

// Evaluation with the Jacobian (Problem.evaluate with compute_jacobian=True,
// problem.py:377-413, 486-530): one thread per observation writes its
// residual block's corrected residuals and Jacobian rows straight into the
// caller's dense matrix - rows 2 row_of_obs[o] and 2 row_of_obs[o] + 1,
// columns from the per-coordinate column maps (-1: no column, a constant
// block or coordinate).  Block partials of the cost.
__global__ __launch_bounds__(kPointThreads) void k_evaluate_jacobian(BAView v, const double* __restrict__ cams,
                                                                      const double* __restrict__ pts,
                                                                      const int32_t* __restrict__ row_of_obs,
                                                                      const int64_t* __restrict__ cam_cols,
                                                                      const int64_t* __restrict__ pt_cols,
                                                                      int64_t ncols, int apply_loss,
                                                                      double* __restrict__ jac,
                                                                      double* __restrict__ residuals,
                                                                      double* __restrict__ part) {
  __shared__ double red[kPointThreads / kWave];
  const int o = blockIdx.x * kPointThreads + threadIdx.x;
  double cost = 0.0;
  if (o < v.N) {
    const int ci = v.obs_camera[o], p = v.obs_point[o];
    const double X[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    const double obs[2] = {v.obs_xy[2 * o], v.obs_xy[2 * o + 1]};
    double r[2], Jc[2][9], Jp[2][3];
    camera_jacobian(v, cams, ci, X, obs, r, Jc, Jp);
    if (apply_loss) {
      cost = robustify2(v.loss, r, Jc, Jp);
    } else {
      cost = 0.5 * (r[0] * r[0] + r[1] * r[1]);
    }
    const size_t row = 2 * (size_t)row_of_obs[o];
    residuals[row] = r[0];
    residuals[row + 1] = r[1];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      double* out = jac + (row + k) * (size_t)ncols;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const int64_t c = cam_cols[9 * ci + j];
        if (c >= 0) out[c] = Jc[k][j];
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int64_t c = pt_cols[3 * p + j];
        if (c >= 0) out[c] = Jp[k][j];
      }
    }
  }
  const double tc = block_sum<kPointThreads>(cost, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tc;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
inline int blocks_for(int n) { return (n + kPointThreads - 1) / kPointThreads; }
}  // namespace

int launch_fill_lds(double value, hipStream_t s) {
  int dev = 0, cus = 0, lds = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  CT_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (lds <= 0 || cus <= 0) return CT_OK;
  CT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fill_lds), hipFuncAttributeMaxDynamicSharedMemorySize,
                             lds));
  hipLaunchKernelGGL(k_fill_lds, dim3(2 * cus), dim3(256), (size_t)lds, s, value, lds / (int)sizeof(double));
  CT_LAUNCH_CHECK();
  return CT_OK;
}

template <typename real>
int launch_eval_obs(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s) {
  if (v.N > 0) {
    hipLaunchKernelGGL(k_eval_obs<real>, dim3(blocks_for(v.N)), dim3(kPointThreads), 0, s, v, b, w);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_eval_fused(const BAView& v, const BABuffers<real>& b, const BAWork& w, int num_groups,
                      bool unobserved_points, hipStream_t s) {
  if (num_groups > 0) {
    const int waves = kFusedThreads / kWave;
    hipLaunchKernelGGL(k_eval_fused<real>, dim3((num_groups + waves - 1) / waves), dim3(kFusedThreads), 0, s, v, b, w,
                       num_groups, unobserved_points ? 1 : 0);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_eval_points(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s) {
  if (v.P > 0) {
    hipLaunchKernelGGL(k_eval_points<real>, dim3(blocks_for(v.P)), dim3(kPointThreads), 0, s, v, b, w);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_camera_items(const BAView& v, const BABuffers<real>& b, const BAWork& w, int num_items, hipStream_t s) {
  if (num_items > 0) {
    hipLaunchKernelGGL(k_camera_items<real>, dim3(num_items), dim3(kWave), 0, s, v, b, w);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_camera_reduce(const BAView& v, const BABuffers<real>& b, const BAWork& w, bool diagonal_only,
                         hipStream_t s) {
  if (v.B > 0) {
    const long long blocks = diagonal_only ? (long long)v.C : (long long)v.C * v.C;
    hipLaunchKernelGGL(k_camera_reduce<real>, dim3((unsigned)((blocks + kReduceBlocks - 1) / kReduceBlocks)),
                       dim3(kReduceThreads), 0, s, v, b, w, diagonal_only ? 1 : 0);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_step(const BAView& v, const BABuffers<real>& b, const BAWork& w, bool fused_backsub, hipStream_t s) {
  const int work = v.P > 9 * v.C ? v.P : 9 * v.C;
  if (work > 0) {
    if (fused_backsub) {
      hipLaunchKernelGGL((k_step_points<real, true>), dim3(blocks_for(work)), dim3(kPointThreads), 0, s, v, b, w);
    } else {
      hipLaunchKernelGGL((k_step_points<real, false>), dim3(blocks_for(work)), dim3(kPointThreads), 0, s, v, b, w);
    }
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_step_candidate(const BAView& v, const BABuffers<real>& b, const BAWork& w, int num_groups, bool backsub,
                          hipStream_t s) {
  const int waves = kFusedThreads / kWave;
  const int grid = std::max((num_groups + waves - 1) / waves, 1);
  if (backsub)
    hipLaunchKernelGGL((k_step_candidate<real, true>), dim3(grid), dim3(kFusedThreads), 0, s, v, b, w, num_groups);
  else
    hipLaunchKernelGGL((k_step_candidate<real, false>), dim3(grid), dim3(kFusedThreads), 0, s, v, b, w, num_groups);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

template <typename real>
int launch_candidate(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s) {
  if (v.N > 0) {
    hipLaunchKernelGGL(k_candidate_obs<real>, dim3(blocks_for(v.N)), dim3(kPointThreads), 0, s, v, b, w);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

template <typename real>
int launch_backsub(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s) {
  if (v.P > 0) {
    hipLaunchKernelGGL(k_backsub<real>, dim3(blocks_for(v.P)), dim3(kPointThreads), 0, s, v, b, w);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

// DOGLEG: gn's point part by back-substitution, the dot products, the
// coefficients and the combined step in dxc / dxp.
template <typename real>
int launch_dogleg_step(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s) {
  CT_TRY(launch_backsub<real>(v, b, w, s));
  const int blocks = dogleg_blocks(v.P);
  hipLaunchKernelGGL(k_dogleg_partials<real>, dim3(blocks), dim3(kPointThreads), 0, s, v, b, w);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dogleg_coefficients, dim3(1), dim3(kDoglegThreads), 0, s, w, blocks);
  CT_LAUNCH_CHECK();
  const size_t work = std::max<size_t>(9 * (size_t)v.C, 3 * (size_t)v.P);
  const unsigned int grid = (unsigned int)std::min<size_t>(std::max<size_t>((work + kPointThreads - 1) / kPointThreads, 1), 4096);
  hipLaunchKernelGGL(k_dogleg_combine<real>, dim3(grid), dim3(kPointThreads), 0, s, v, b, w);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_poison_step(const BAView& v, const BAWork& w, hipStream_t s) {
  const size_t work = std::max<size_t>(std::max<size_t>(9 * (size_t)v.C, 3 * (size_t)v.P), 1);
  const unsigned int grid = (unsigned int)std::min<size_t>((work + kPointThreads - 1) / kPointThreads, 1024);
  hipLaunchKernelGGL(k_poison_step, dim3(grid), dim3(kPointThreads), 0, s, w, 9 * v.C, 3 * v.P);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

template <typename real>
int launch_refinement(const BAView& v, const BABuffers<real>& b, const BAWork& w, int steps, hipStream_t s) {
  if (steps <= 0) return launch_backsub<real>(v, b, w, s);
  for (int k = 0; k < steps; ++k) {
    if (v.P > 0) {
      if (k == 0) hipLaunchKernelGGL((k_refine_pass<real, true>), dim3(blocks_for(v.P)), dim3(kPointThreads), 0, s, v, b, w);
      else hipLaunchKernelGGL((k_refine_pass<real, false>), dim3(blocks_for(v.P)), dim3(kPointThreads), 0, s, v, b, w);
      CT_LAUNCH_CHECK();
    }
    if (v.C > 0) {
      if (k == 0) hipLaunchKernelGGL((k_refine_cameras<real, false>), dim3(v.C), dim3(kWave), 0, s, v, b, w);
      else hipLaunchKernelGGL((k_refine_cameras<real, true>), dim3(v.C), dim3(kWave), 0, s, v, b, w);
      CT_LAUNCH_CHECK();
      CT_TRY(launch_rcs_tri_solve(w.S, w.res_c, w.dxc_corr, 9 * v.C, w.state, w.solve_tmp, s));
    }
  }
  const int work = v.P > 9 * v.C ? v.P : 9 * v.C;
  if (work > 0) {
    hipLaunchKernelGGL(k_refine_update<real>, dim3(blocks_for(work)), dim3(kPointThreads), 0, s, v, b, w);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

int launch_evaluate_only(const BAView& v, const double* cams, const double* pts, double* residuals, double* part,
                         hipStream_t s) {
  if (v.P > 0) {
    hipLaunchKernelGGL(k_evaluate_only, dim3(blocks_for(v.P)), dim3(kPointThreads), 0, s, v, cams, pts, residuals,
                       part);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

int launch_evaluate_jacobian(const BAView& v, const double* cams, const double* pts, const int32_t* row_of_obs,
                             const int64_t* cam_cols, const int64_t* pt_cols, int64_t ncols, int apply_loss,
                             double* jac, double* residuals, double* part, hipStream_t s) {
  if (v.N > 0) {
    hipLaunchKernelGGL(k_evaluate_jacobian, dim3(blocks_for(v.N)), dim3(kPointThreads), 0, s, v, cams, pts, row_of_obs,
                       cam_cols, pt_cols, ncols, apply_loss, jac, residuals, part);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

#define CT_INSTANTIATE(real)                                                                           \
  template int launch_eval_obs<real>(const BAView&, const BABuffers<real>&, const BAWork&, hipStream_t);        \
  template int launch_eval_points<real>(const BAView&, const BABuffers<real>&, const BAWork&, hipStream_t);     \
  template int launch_camera_items<real>(const BAView&, const BABuffers<real>&, const BAWork&, int, hipStream_t); \
  template int launch_camera_reduce<real>(const BAView&, const BABuffers<real>&, const BAWork&, bool, hipStream_t); \
  template int launch_step<real>(const BAView&, const BABuffers<real>&, const BAWork&, bool, hipStream_t);      \
  template int launch_candidate<real>(const BAView&, const BABuffers<real>&, const BAWork&, hipStream_t);
CT_INSTANTIATE(float)
CT_INSTANTIATE(double)
#undef CT_INSTANTIATE
template int launch_step_candidate<float>(const BAView&, const BABuffers<float>&, const BAWork&, int, bool,
                                         hipStream_t);
template int launch_step_candidate<double>(const BAView&, const BABuffers<double>&, const BAWork&, int, bool,
                                          hipStream_t);
template int launch_eval_fused<float>(const BAView&, const BABuffers<float>&, const BAWork&, int, bool,
                                       hipStream_t);
template int launch_eval_fused<double>(const BAView&, const BABuffers<double>&, const BAWork&, int, bool,
                                        hipStream_t);
template int launch_backsub<float>(const BAView&, const BABuffers<float>&, const BAWork&, hipStream_t);
template int launch_backsub<double>(const BAView&, const BABuffers<double>&, const BAWork&, hipStream_t);
template int launch_dogleg_step<float>(const BAView&, const BABuffers<float>&, const BAWork&, hipStream_t);
template int launch_dogleg_step<double>(const BAView&, const BABuffers<double>&, const BAWork&, hipStream_t);
template int launch_refinement<float>(const BAView&, const BABuffers<float>&, const BAWork&, int, hipStream_t);
template int launch_refinement<double>(const BAView&, const BABuffers<double>&, const BAWork&, int, hipStream_t);

}  // namespace ct
