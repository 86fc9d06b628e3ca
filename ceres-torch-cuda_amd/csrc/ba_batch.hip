// Batched small bundle adjustments: one persistent launch, one workgroup per
// problem.  A workgroup takes a problem from a work queue and runs its whole
// LM solve (solver.py:35-357) with the DENSE_SCHUR step (schur_solve_dense,
// linear.py:667-699) before it takes the next, so uneven problems balance
// over the compute units by themselves.  Per LM iteration, with barriers
// between the phases:
//   E1  per observation: residual, Jacobian, loss correction (Jc, Jp, r);
//       skipped when x is unchanged since the last evaluation (a rejected
//       step changes only the damping);
//   E2  per point: V_p = sum Jp^T Jp + D_p, g_p, the 3x3 factor L, q = L^-1 g_p,
//       P_o = Jp_o L^-T and e_o = P_o q - r_o per observation;
//   A1  per work item (at most kItemWork observation pairs of one camera-pair
//       block i <= j, or kItemWork observations of one camera) and 3-row
//       strip: the item's partial of S = U + D_c - sum Jc_a^T (P_a P_b^T) Jc_b
//       and, for observation items, of the column norms, g_c and sum Jc^T e;
//   A2  per element of S and per camera column: the item partials summed in
//       item order, the damping D_c, S written, rhs, g_c;
//   (the gradient norms of E2 and A2 go to the LM begin update)
//   C   S factored with the right-hand side as an extra row by the
//       one-workgroup Cholesky (16x16 tiles on v_mfma_f64_16x16x4f64, the
//       packed factor in LDS), then the backward substitution;
//   B   point back-substitution dx_p = -L^-T (q + sum P^T Jc dx_c), Plus,
//       candidate cost and model decrease -r.u - u.u/2 per observation.
// The LM state lives in LDS and thread 0 runs the trust-region arithmetic of
// lm_logic.h on it.  Every sum has a fixed order (per-thread strided sums,
// then wave trees and waves in index order), so a problem's result does not
// depend on the rest of the batch or on which workgroup solved it.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ba_kernels.h"
#include "ba_camera.h"
#include "chol_small.h"
#include "lm_logic.h"

namespace ct {

namespace {

constexpr int kBT = kSmallThreads;  // 8 waves: the one-workgroup Cholesky's shape
constexpr int kBWaves = kBT / kWave;
static_assert(9 * CT_BA_BATCH_MAX_CAMERAS <= kSmallMax, "the batched reduced camera system exceeds the small Cholesky");

// Per-problem workspace slice (doubles, each array 256-byte aligned).
struct BatchWork {
  double* xc;    // [2][9C] ping-pong camera parameters
  double* xp;    // [2][3P] ping-pong point parameters
  double* r;     // [N][2] corrected residuals at x
  double* Jp;    // [N][6] corrected point Jacobians at x
  double* JP;    // [N][24] corrected camera Jacobian (2x9), then P = Jp L^-T (2x3)
  double* e;     // [N][2] P q - r
  double* Linv;  // [P][6] L^-1, lower triangle row-major
  double* q;     // [P][3]
  double* dxp;   // [P][3]
  double* S;     // [n][n] reduced camera system (lower triangle used)
  double* rhs;   // [n]
  double* dxc;   // [n]
  double* itemS;    // [items][81] partial camera-pair blocks
  double* itemV;    // [observation items][27] column norms, g_c and sum Jc^T e of the item's columns
};

constexpr int kItemWork = 32;  // observation pairs or observations per work item
constexpr int kMaxBlocks = CT_BA_BATCH_MAX_CAMERAS * (CT_BA_BATCH_MAX_CAMERAS + 1) / 2;

// Work items of a problem with B camera-pair blocks, T pairs and N observations (an upper bound:
// each block's or camera's last item may be partial).
__host__ __device__ inline size_t max_items(int C, int64_t T, int N) {
  const size_t B = (size_t)C * (C + 1) / 2;
  return (size_t)(T / kItemWork) + B + (size_t)(N / kItemWork) + (size_t)C + 1;
}

__host__ __device__ inline size_t batch_layout(int C, int P, int N, int64_t T, double* base, BatchWork* w) {
  const size_t n = 9 * (size_t)C;
  const size_t items = max_items(C, T, N);
  const size_t obs_items = (size_t)(N / kItemWork) + (size_t)C + 1;
  size_t off = 0;
  double** slots[] = {&w->xc, &w->xp, &w->r, &w->Jp, &w->JP, &w->e, &w->Linv, &w->q, &w->dxp, &w->S, &w->rhs,
                      &w->dxc, &w->itemS, &w->itemV};
  const size_t counts[] = {2 * n, 6 * (size_t)P, 2 * (size_t)N, 6 * (size_t)N, 24 * (size_t)N, 2 * (size_t)N,
                           6 * (size_t)P, 3 * (size_t)P, 3 * (size_t)P, n * n, n, n, 81 * items, 27 * obs_items};
  for (int k = 0; k < 14; ++k) {
    off = (off + 31) & ~size_t(31);
    *slots[k] = base ? base + off : nullptr;
    off += counts[k];
  }
  return ((off * sizeof(double)) + 255) & ~size_t(255);
}

struct BatchArgs {
  ct_ba_batch b;
  ct_lm_options opt;
  double* cameras;
  double* points;
  ct_lm_report* results;
  double* records;
  char* work;
  size_t work_bytes;
  unsigned int* queue;
};

// Every wave's own memory operations have completed, then the workgroup
// barrier: values written to global memory by one wave are read by others
// after it (all waves of a workgroup share one compute unit's L1).
// A value every lane holds alike (read from LDS or through a loop), made
// known to be uniform so that what derives from it stays in scalar registers.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// threadIdx.x as an opaque value, read afresh by each phase: the compiler may
// not hoist what a phase derives from it out of the LM loop, where it would
// stay live across the other phases.
__device__ __forceinline__ int fresh_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

__device__ __forceinline__ void wg_sync() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

// Fixed-order sums of K values over the workgroup (wave trees, then waves in
// index order); the totals are valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double (*scratch)[K]) {
  const int ft = fresh_tid();
  const int lane = ft & (kWave - 1), wid = ft / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[wid][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int w = 0; w < kBWaves; ++w) t += scratch[w][k];
      v[k] = t;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void load_record(const double* __restrict__ src, double (&v)[24]) {
  const double2* s2 = reinterpret_cast<const double2*>(src);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double2 t = s2[k];
    v[2 * k] = t.x;
    v[2 * k + 1] = t.y;
  }
}

// Camera pair (i, j), i <= j, of block b in row-major order over the upper triangle.
__device__ __forceinline__ void block_pair(int b, int C, int& i, int& j) {
  int r = b, row = 0;
  while (r >= C - row) {
    r -= C - row;
    ++row;
  }
  i = row;
  j = row + r;
}

struct Problem {
  int C, P, N, n;
  const int32_t* oc;    // obs_camera
  const int32_t* op;    // obs_point
  const double* xy;     // observations
  const int32_t* poff;  // point_offsets
  const int32_t* coff;  // camera_offsets
  const int32_t* cobs;  // camera_obs
  const int32_t* boff;  // block_offsets
  const int32_t* terms; // block_terms
  ct_loss loss;
};

// E1: residual, Jacobian and loss correction per observation at x; returns
// this thread's part of the cost.
__device__ double eval_observations(const Problem& pb, const BatchWork& w, const double* __restrict__ xc,
                                    const double* __restrict__ xp) {
  double cost = 0.0;
  for (int o = fresh_tid(); o < pb.N; o += kBT) {
    const int ci = pb.oc[o], p = pb.op[o];
    const double X[3] = {xp[3 * p], xp[3 * p + 1], xp[3 * p + 2]};
    const double obs[2] = {pb.xy[2 * o], pb.xy[2 * o + 1]};
    double cam[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) cam[j] = xc[9 * ci + j];
    double r[2], Jc[2][9], Jp[2][3];
    snavely_jacobian(cam, X, obs, r, Jc, Jp);
    cost += robustify2(pb.loss, r, Jc, Jp);
    w.r[2 * (size_t)o] = r[0];
    w.r[2 * (size_t)o + 1] = r[1];
    double* jp = w.Jp + 6 * (size_t)o;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j) jp[3 * k + j] = Jp[k][j];
    double* rec = w.JP + 24 * (size_t)o;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 9; ++j) rec[9 * k + j] = Jc[k][j];
  }
  return cost;
}

// E2: point elimination (the arithmetic of k_eval_points in FP64).
__device__ void eliminate_points(const Problem& pb, const BatchWork& w, const LMState* st, double& g2,
                                 double& gmax) {
  for (int p = fresh_tid(); p < pb.P; p += kBT) {
    const int o0 = pb.poff[p], o1 = pb.poff[p + 1];
    double* Li = w.Linv + 6 * (size_t)p;
    if (o0 == o1) {
      // no observations: the block is the damping alone, the gradient and step zero
      double inv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double d = column_damping(st, 0.0, true);
        inv[k] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
        w.q[3 * p + k] = 0.0;
      }
      Li[0] = inv[0]; Li[1] = 0.0; Li[2] = inv[1]; Li[3] = 0.0; Li[4] = 0.0; Li[5] = inv[2];
      continue;
    }
    double V00 = 0, V10 = 0, V11 = 0, V20 = 0, V21 = 0, V22 = 0;
    double g0 = 0.0, g1 = 0.0, g2v = 0.0;
    for (int o = o0; o < o1; ++o) {
      const double* jp = w.Jp + 6 * (size_t)o;
      const double a0 = jp[0], a1 = jp[1], a2 = jp[2], c0 = jp[3], c1 = jp[4], c2 = jp[5];
      const double r0 = w.r[2 * (size_t)o], r1 = w.r[2 * (size_t)o + 1];
      V00 += a0 * a0 + c0 * c0;
      V10 += a1 * a0 + c1 * c0;
      V11 += a1 * a1 + c1 * c1;
      V20 += a2 * a0 + c2 * c0;
      V21 += a2 * a1 + c2 * c1;
      V22 += a2 * a2 + c2 * c2;
      g0 += a0 * r0 + c0 * r1;
      g1 += a1 * r0 + c1 * r1;
      g2v += a2 * r0 + c2 * r1;
    }
    // damping D = clamp(colsum J^2)/radius (linear.py:539-551)
    V00 += column_damping(st, V00, true);
    V11 += column_damping(st, V11, true);
    V22 += column_damping(st, V22, true);
    const double l00 = sqrt(V00);
    const double l10 = V10 / l00, l20 = V20 / l00;
    const double l11 = sqrt(V11 - l10 * l10);
    const double l21 = (V21 - l20 * l10) / l11;
    const double l22 = sqrt(V22 - l20 * l20 - l21 * l21);
    const double i00 = 1.0 / l00, i11 = 1.0 / l11, i22 = 1.0 / l22;
    const double i10 = -l10 * i00 * i11;
    const double i21 = -l21 * i11 * i22;
    const double i20 = -(l20 * i00 + l21 * i10) * i22;
    const double q0 = i00 * g0;
    const double q1 = i10 * g0 + i11 * g1;
    const double q2 = i20 * g0 + i21 * g1 + i22 * g2v;
    Li[0] = i00; Li[1] = i10; Li[2] = i11; Li[3] = i20; Li[4] = i21; Li[5] = i22;
    w.q[3 * p] = q0;
    w.q[3 * p + 1] = q1;
    w.q[3 * p + 2] = q2;
    for (int o = o0; o < o1; ++o) {
      const double* jp = w.Jp + 6 * (size_t)o;
      double* P = w.JP + 24 * (size_t)o + 18;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double j0 = jp[3 * k], j1 = jp[3 * k + 1], j2 = jp[3 * k + 2];
        // P = Jp L^-T : P[k][m] = sum_{l<=m} Jp[k][l] Linv[m][l]
        const double p0 = j0 * i00;
        const double p1 = j0 * i10 + j1 * i11;
        const double p2 = j0 * i20 + j1 * i21 + j2 * i22;
        P[3 * k] = p0;
        P[3 * k + 1] = p1;
        P[3 * k + 2] = p2;
        w.e[2 * (size_t)o + k] = (p0 * q0 + p1 * q1 + p2 * q2) - w.r[2 * (size_t)o + k];
      }
    }
    g2 += g0 * g0 + g1 * g1 + g2v * g2v;
    gmax = fmax(gmax, fmax(fabs(g0), fmax(fabs(g1), fabs(g2v))));
  }
}

// Item index of a problem (LDS, built once per problem): the first item of
// each camera-pair block, then the first observation item of each camera.
struct ItemIndex {
  int blk0[kMaxBlocks + 1];               // pair items of block b: [blk0[b], blk0[b + 1])
  int cam0[CT_BA_BATCH_MAX_CAMERAS + 1];  // observation items of camera i: [cam0[i], cam0[i + 1])
};

__device__ void build_item_index(const Problem& pb, ItemIndex& ix) {
  if (threadIdx.x != 0) return;
  const int B = pb.C * (pb.C + 1) / 2;
  ix.blk0[0] = 0;
  for (int b = 0; b < B; ++b) ix.blk0[b + 1] = ix.blk0[b] + (pb.boff[b + 1] - pb.boff[b] + kItemWork - 1) / kItemWork;
  ix.cam0[0] = ix.blk0[B];
  for (int i = 0; i < pb.C; ++i)
    ix.cam0[i + 1] = ix.cam0[i] + (pb.coff[i + 1] - pb.coff[i] + kItemWork - 1) / kItemWork;
}

// Last k in [0, count) with a[k] <= v (a non-decreasing, a[0] <= v).
__device__ __forceinline__ int last_le(const int* a, int count, int v) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A1: one unit per (item, 3-row strip s).  A pair item sums
// -Jc_a^T (P_a P_b^T) Jc_b over its pairs; an observation item of camera i
// sums U = Jc^T Jc and, for the strip's three columns, the column norm,
// g_c = Jc^T r and Jc^T e.  The strip's rows of Jc_a are read by offset (a
// register array indexed by the runtime strip would live in scratch memory).
__device__ void camera_items(const Problem& pb, const BatchWork& w, const ItemIndex& ix) {
  const int B = pb.C * (pb.C + 1) / 2;
  const int I0 = ix.blk0[B], I = ix.cam0[pb.C];
  for (int u = fresh_tid(); u < 3 * I; u += kBT) {
    const int q = u / 3, s = u - 3 * q;
    double acc[3][9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[r][c] = 0.0;
    if (q < I0) {
      const int b = last_le(ix.blk0, B, q);
      const int t0 = pb.boff[b] + (q - ix.blk0[b]) * kItemWork;
      const int t1 = min(t0 + kItemWork, pb.boff[b + 1]);
      for (int t = t0; t < t1; ++t) {
        const int oa = pb.terms[2 * (size_t)t], ob = pb.terms[2 * (size_t)t + 1];
        const double* ra = w.JP + 24 * (size_t)oa;
        double rb[24], pa[6], a0[3], a1[3];
        load_record(w.JP + 24 * (size_t)ob, rb);
#pragma unroll
        for (int k = 0; k < 6; ++k) pa[k] = ra[18 + k];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          a0[r] = ra[3 * s + r];
          a1[r] = ra[9 + 3 * s + r];
        }
        // M = P_a P_b^T (2x2), T = M Jc_b (2x9)
        const double m00 = pa[0] * rb[18] + pa[1] * rb[19] + pa[2] * rb[20];
        const double m01 = pa[0] * rb[21] + pa[1] * rb[22] + pa[2] * rb[23];
        const double m10 = pa[3] * rb[18] + pa[4] * rb[19] + pa[5] * rb[20];
        const double m11 = pa[3] * rb[21] + pa[4] * rb[22] + pa[5] * rb[23];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
          const double t0v = m00 * rb[c] + m01 * rb[9 + c];
          const double t1v = m10 * rb[c] + m11 * rb[9 + c];
#pragma unroll
          for (int r = 0; r < 3; ++r) acc[r][c] -= a0[r] * t0v + a1[r] * t1v;
        }
      }
    } else {
      const int i = last_le(ix.cam0, pb.C, q);
      const int t0 = pb.coff[i] + (q - ix.cam0[i]) * kItemWork;
      const int t1 = min(t0 + kItemWork, pb.coff[i + 1]);
      double colsq[3] = {0.0, 0.0, 0.0}, gc[3] = {0.0, 0.0, 0.0}, rh[3] = {0.0, 0.0, 0.0};
      for (int t = t0; t < t1; ++t) {
        const int o = pb.cobs[t];
        const double* ro = w.JP + 24 * (size_t)o;
        double jc[18], a0[3], a1[3];
#pragma unroll
        for (int k = 0; k < 18; ++k) jc[k] = ro[k];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          a0[r] = ro[3 * s + r];
          a1[r] = ro[9 + 3 * s + r];
        }
        const double r0 = w.r[2 * (size_t)o], r1 = w.r[2 * (size_t)o + 1];
        const double e0 = w.e[2 * (size_t)o], e1 = w.e[2 * (size_t)o + 1];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
          for (int c = 0; c < 9; ++c) acc[r][c] += a0[r] * jc[c] + a1[r] * jc[9 + c];
          colsq[r] += a0[r] * a0[r] + a1[r] * a1[r];
          gc[r] += a0[r] * r0 + a1[r] * r1;
          rh[r] += a0[r] * e0 + a1[r] * e1;
        }
      }
      double* v = w.itemV + 27 * (size_t)(q - I0);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        v[3 * s + r] = colsq[r];
        v[9 + 3 * s + r] = gc[r];
        v[18 + 3 * s + r] = rh[r];
      }
    }
    double* dst = w.itemS + 81 * (size_t)q + 27 * s;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 9; ++c) dst[9 * r + c] = acc[r][c];
  }
}

// A2: one unit per element of S's camera-pair blocks (row-major within the
// block) and one per camera column, summing the item partials in item order.
// A block off the diagonal is stored transposed below the diagonal (the
// Cholesky reads the lower triangle); a diagonal block whole, with the
// damping of its columns (linear.py:539-551) from the observation items'
// column norms, summed by the diagonal element in the same order as by its
// column unit.
__device__ void reduce_rcs(const Problem& pb, const BatchWork& w, const ItemIndex& ix, const LMState* st,
                           double& g2, double& gmax) {
  const int B = pb.C * (pb.C + 1) / 2, n = pb.n;
  const int I0 = ix.blk0[B];
  for (int u = fresh_tid(); u < 81 * B + n; u += kBT) {
    if (u < 81 * B) {
      const int b = u / 81, e = u - 81 * b;
      const int r = e / 9, c = e - 9 * r;
      int i, j;
      block_pair(b, pb.C, i, j);
      double sum = 0.0;
      for (int q = ix.blk0[b]; q < ix.blk0[b + 1]; ++q) sum += w.itemS[81 * (size_t)q + e];
      if (i == j) {
        double colsq = 0.0;
        for (int q = ix.cam0[i]; q < ix.cam0[i + 1]; ++q) {
          sum += w.itemS[81 * (size_t)q + e];
          colsq += w.itemV[27 * (size_t)(q - I0) + c];
        }
        if (r == c) sum += column_damping(st, colsq, true);
        w.S[(size_t)(9 * i + r) * n + 9 * i + c] = sum;
      } else {
        w.S[(size_t)(9 * j + c) * n + 9 * i + r] = sum;
      }
    } else {
      const int k = u - 81 * B, i = k / 9, c = k - 9 * i;
      double gc = 0.0, rh = 0.0;
      for (int q = ix.cam0[i]; q < ix.cam0[i + 1]; ++q) {
        const double* v = w.itemV + 27 * (size_t)(q - I0);
        gc += v[9 + c];
        rh += v[18 + c];
      }
      w.rhs[k] = rh;
      g2 += gc * gc;
      gmax = fmax(gmax, fabs(gc));
    }
  }
}

// B1: camera Plus and the point back-substitution, written to the other
// parameter copy; |dx|^2, |x'|^2 and a non-finite count.
__device__ void step_parameters(const Problem& pb, const BatchWork& w, int cur, double (&v)[5]) {
  const double* xc = w.xc + (size_t)cur * pb.n;
  const double* xp = w.xp + (size_t)cur * 3 * pb.P;
  double* nxc = w.xc + (size_t)(1 - cur) * pb.n;
  double* nxp = w.xp + (size_t)(1 - cur) * 3 * pb.P;
  for (int t = fresh_tid(); t < pb.n; t += kBT) {
    const double d = w.dxc[t];
    const double nx = xc[t] + d;
    nxc[t] = nx;
    v[0] += d * d;
    v[1] += nx * nx;
    if (!isfinite(d)) v[2] = 1.0;
  }
  for (int p = fresh_tid(); p < pb.P; p += kBT) {
    // dx_p = -L^-T (q + sum_o P_o^T Jc_o dx_c)   (schur_solve_dense's xa, linear.py:698)
    double s0 = w.q[3 * p], s1 = w.q[3 * p + 1], s2 = w.q[3 * p + 2];
    for (int o = pb.poff[p]; o < pb.poff[p + 1]; ++o) {
      const int ci = pb.oc[o];
      double rec[24];
      load_record(w.JP + 24 * (size_t)o, rec);
      double y0 = 0.0, y1 = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double d = w.dxc[9 * ci + j];
        y0 += rec[j] * d;
        y1 += rec[9 + j] * d;
      }
      s0 += rec[18] * y0 + rec[21] * y1;
      s1 += rec[19] * y0 + rec[22] * y1;
      s2 += rec[20] * y0 + rec[23] * y1;
    }
    const double* Li = w.Linv + 6 * (size_t)p;
    const double d0 = -(Li[0] * s0 + Li[1] * s1 + Li[3] * s2);
    const double d1 = -(Li[2] * s1 + Li[4] * s2);
    const double d2 = -(Li[5] * s2);
    w.dxp[3 * p] = d0;
    w.dxp[3 * p + 1] = d1;
    w.dxp[3 * p + 2] = d2;
    const double n0 = xp[3 * p] + d0, n1 = xp[3 * p + 1] + d1, n2 = xp[3 * p + 2] + d2;
    nxp[3 * p] = n0;
    nxp[3 * p + 1] = n1;
    nxp[3 * p + 2] = n2;
    v[0] += d0 * d0 + d1 * d1 + d2 * d2;
    v[1] += n0 * n0 + n1 * n1 + n2 * n2;
    if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) v[2] = 1.0;
  }
}

// B2: candidate cost at x' and the model decrease -r.u - u.u/2, u = J dx, per observation.
__device__ void candidate(const Problem& pb, const BatchWork& w, int cur, double (&v)[5]) {
  const double* nxc = w.xc + (size_t)(1 - cur) * pb.n;
  const double* nxp = w.xp + (size_t)(1 - cur) * 3 * pb.P;
  for (int o = fresh_tid(); o < pb.N; o += kBT) {
    const int ci = pb.oc[o], p = pb.op[o];
    const double* jq = w.Jp + 6 * (size_t)o;
    const double* jc = w.JP + 24 * (size_t)o;
    const double d0 = w.dxp[3 * p], d1 = w.dxp[3 * p + 1], d2 = w.dxp[3 * p + 2];
    double u0 = jq[0] * d0 + jq[1] * d1 + jq[2] * d2;
    double u1 = jq[3] * d0 + jq[4] * d1 + jq[5] * d2;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const double d = w.dxc[9 * ci + j];
      u0 += jc[j] * d;
      u1 += jc[9 + j] * d;
    }
    const double r0 = w.r[2 * (size_t)o], r1 = w.r[2 * (size_t)o + 1];
    v[4] += -(r0 * u0 + r1 * u1) - 0.5 * (u0 * u0 + u1 * u1);
    const double X[3] = {nxp[3 * p], nxp[3 * p + 1], nxp[3 * p + 2]};
    const double obs[2] = {pb.xy[2 * o], pb.xy[2 * o + 1]};
    double rr[2];
    snavely_residual(nxc + 9 * (size_t)ci, X, obs, rr);
    v[3] += robust_cost2(pb.loss, rr);
  }
}

struct BatchShared {
  SmallCholShared chol;
  ItemIndex items;
  LMState st;
  double red[kBWaves][5];
  double red3[kBWaves][3];
  int slot;
};

__device__ void report_invalid(const BatchArgs& a, int pi) {
  if (threadIdx.x != 0 || a.results == nullptr) return;
  ct_lm_report rep{};
  rep.termination = CT_TERM_FAILURE;
  rep.reason = CT_REASON_INVALID_INPUT;
  a.results[pi] = rep;
}

__device__ void solve_problem(const BatchArgs& a, int pi, BatchShared& sh) {
  const ct_ba_batch_problem pr = a.b.problems[pi];
  Problem pb;
  pb.C = pr.num_cameras;
  pb.P = pr.num_points;
  pb.N = pr.num_observations;
  pb.n = 9 * pb.C;
  if (pb.C < 1 || pb.C > CT_BA_BATCH_MAX_CAMERAS || pb.P < 0 || pb.N < 0 || pr.num_terms < 0 ||
      (pr.work_base & 255) != 0 || pr.work_base < 256) {
    report_invalid(a, pi);
    return;
  }
  BatchWork w;
  const size_t slice = batch_layout(pb.C, pb.P, pb.N, pr.num_terms, nullptr, &w);
  if ((size_t)pr.work_base + slice > a.work_bytes) {
    report_invalid(a, pi);
    return;
  }
  pb.oc = a.b.obs_camera + pr.obs_base;
  pb.op = a.b.obs_point + pr.obs_base;
  pb.xy = a.b.observations + 2 * pr.obs_base;
  pb.poff = a.b.point_offsets + pr.point_offsets_base;
  pb.coff = a.b.camera_offsets + pr.camera_offsets_base;
  pb.cobs = a.b.camera_obs + pr.obs_base;
  pb.boff = a.b.block_offsets + pr.block_offsets_base;
  pb.terms = a.b.block_terms + 2 * pr.term_base;
  pb.loss = pr.loss;
  batch_layout(pb.C, pb.P, pb.N, pr.num_terms, reinterpret_cast<double*>(a.work + pr.work_base), &w);
  double* ucam = a.cameras + 9 * pr.camera_base;
  double* upt = a.points + 3 * pr.point_base;
  double* rec = a.records ? a.records + (size_t)pi * (a.opt.max_num_iterations + 2) * CT_REC_FIELDS : nullptr;
  LMState* st = &sh.st;
  const int tid = threadIdx.x;

  for (int i = tid; i < pb.n; i += kBT) {
    const double v = ucam[i];
    w.xc[i] = v;
    w.xc[pb.n + i] = v;
  }
  for (int i = tid; i < 3 * pb.P; i += kBT) {
    const double v = upt[i];
    w.xp[i] = v;
    w.xp[3 * pb.P + i] = v;
  }
  if (tid == 0) lm_state_init(st, a.opt);
  build_item_index(pb, sh.items);
  wg_sync();

  // thread 0 keeps the evaluation at x for the passes that reuse it
  double cost_x = 0.0, g2_x = 0.0, gmax_x = 0.0;
  int jac_cur = -1;    // the parameter copy whose residuals and Jacobians are stored
  bool reuse = false;  // x and the radius are unchanged since the last evaluation
  const int max_passes = a.opt.max_num_iterations + 2;
  for (int pass = 0; pass < max_passes; ++pass) {
    const int cur = uniform(st->cur);
    if (!reuse) {
      double part[3] = {0.0, 0.0, 0.0};  // cost, |g|^2 (max separately)
      double gmax = 0.0;
      const bool jac = jac_cur != cur;
      if (jac) part[0] = eval_observations(pb, w, w.xc + (size_t)cur * pb.n, w.xp + (size_t)cur * 3 * pb.P);
      wg_sync();
      eliminate_points(pb, w, st, part[1], gmax);
      wg_sync();
      camera_items(pb, w, sh.items);
      wg_sync();
      reduce_rcs(pb, w, sh.items, st, part[1], gmax);
      block_sums<3>(part, sh.red3);
      const double gm = block_max<kBT>(gmax, &sh.red[0][0]);
      if (tid == 0) {
        if (jac) cost_x = part[0];
        g2_x = part[1];
        gmax_x = gm;
      }
      jac_cur = cur;
    }
    if (tid == 0) lm_begin_update(st, cost_x, g2_x, gmax_x, rec);
    wg_sync();
    if (uniform(st->done)) break;
    if (uniform(st->skip_step)) {
      wg_sync();  // every thread has read the flag before thread 0 clears it
      if (tid == 0) lm_skip_update(st);
      reuse = true;
      wg_sync();
      continue;
    }
    reuse = false;
    small_factor<9 * CT_BA_BATCH_MAX_CAMERAS, false>(w.S, w.rhs, pb.n, sh.chol);
    double* y = sh.chol.Lp + packed_index(pb.n, 0);  // row n: y = L^-1 rhs
    small_backward(sh.chol.Lp, sh.chol.dinv, y, pb.n);
    int my_bad = 0;
    for (int i = tid; i < pb.n; i += kBT) {
      w.dxc[i] = y[i];
      if (!isfinite(y[i])) my_bad = 1;
    }
    if (my_bad) sh.chol.bad = 1;
    wg_sync();
    if (tid == 0) {
      if (sh.chol.bad) st->step_invalid = 1;
      st->linear_iterations = 1;
    }
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // |dx|^2, |x'|^2, non-finite, candidate cost, model decrease
    step_parameters(pb, w, cur, v);
    wg_sync();
    candidate(pb, w, cur, v);
    block_sums<5>(v, sh.red);
    if (tid == 0) lm_end_update(st, v[3], v[4], v[0], v[1], v[2] != 0.0, rec);
    wg_sync();
    if (uniform(st->done)) break;
  }
  const int cur = uniform(st->cur);
  for (int i = tid; i < pb.n; i += kBT) ucam[i] = w.xc[(size_t)cur * pb.n + i];
  for (int i = tid; i < 3 * pb.P; i += kBT) upt[i] = w.xp[(size_t)cur * 3 * pb.P + i];
  if (tid == 0 && a.results) {
    ct_lm_report rep{};
    rep.termination = st->done ? st->termination : CT_TERM_FAILURE;
    rep.reason = st->done ? st->reason : CT_REASON_NONE;
    rep.num_records = st->num_records;
    rep.num_successful_steps = st->num_successful;
    rep.num_unsuccessful_steps = st->num_unsuccessful;
    rep.num_residual_evaluations = st->num_residual_evals;
    rep.num_jacobian_evaluations = st->num_jacobian_evals;
    rep.num_linear_solves = st->num_linear_solves;
    rep.initial_cost = st->initial_cost;
    rep.final_cost = st->best_cost;
    rep.device_seconds = (double)(st->t_end - st->t0) * 1e-8;
    a.results[pi] = rep;
  }
  wg_sync();
}

__global__ __launch_bounds__(kBT) void k_ba_batch(BatchArgs a) {
  __shared__ BatchShared sh;
  for (;;) {
    if (threadIdx.x == 0) sh.slot = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int k = uniform(sh.slot);
    __syncthreads();  // every thread has read the slot before it is taken again
    if (k >= a.b.num_problems) break;
    solve_problem(a, uniform(a.b.order ? a.b.order[k] : k), sh);
  }
}

__global__ void k_ba_batch_reset(unsigned int* q) {
  if (threadIdx.x == 0) q[0] = 0u;
}

}  // namespace

}  // namespace ct

using namespace ct;

extern "C" int ct_ba_batch_problem_bytes(int32_t num_cameras, int32_t num_points, int32_t num_observations,
                                         int64_t num_terms, size_t* bytes) {
  clear_error();
  CT_CHECK_ARG(bytes != nullptr, "bytes is null");
  CT_CHECK_ARG(num_cameras >= 1 && num_cameras <= CT_BA_BATCH_MAX_CAMERAS,
               "a batched bundle adjustment has 1..CT_BA_BATCH_MAX_CAMERAS cameras");
  CT_CHECK_ARG(num_points >= 0 && num_observations >= 0 && num_terms >= 0, "negative problem size");
  BatchWork w;
  *bytes = batch_layout(num_cameras, num_points, num_observations, num_terms, nullptr, &w);
  return CT_OK;
}

extern "C" int ct_ba_batch_solve(const ct_ba_batch* batch, const ct_lm_options* options, double* cameras,
                                 double* points, ct_lm_report* results, double* records, void* workspace,
                                 size_t workspace_bytes, ct_stream_t stream) {
  clear_error();
  CT_CHECK_ARG(batch != nullptr && options != nullptr, "null argument");
  CT_CHECK_ARG(batch->num_problems >= 0, "negative batch size");
  CT_CHECK_ARG(options->max_num_iterations >= 0, "max_num_iterations must be >= 0");
  CT_CHECK_ARG(options->initial_trust_region_radius > 0.0, "initial_trust_region_radius must be > 0");
  if (options->trust_region_strategy != CT_LEVENBERG_MARQUARDT || options->linear_solver != CT_DENSE_SCHUR ||
      options->precision != CT_PLAIN_FP64) {
    set_error("batched bundle adjustment runs LEVENBERG_MARQUARDT with DENSE_SCHUR in plain FP64");
    return CT_ERR_UNSUPPORTED;
  }
  if (options->debug_invalid_steps != 0) {
    set_error("debug_invalid_steps is a test hook of the per-problem bundle-adjustment plans");
    return CT_ERR_UNSUPPORTED;
  }
  if (batch->num_problems == 0) return CT_OK;
  CT_CHECK_ARG(batch->problems && batch->obs_camera && batch->obs_point && batch->observations &&
                   batch->point_offsets && batch->camera_offsets && batch->camera_obs && batch->block_offsets &&
                   batch->block_terms && cameras && points,
               "null batch array");
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= 256, "workspace too small");
  int dev = 0, cus = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  BatchArgs a;
  a.b = *batch;
  a.opt = *options;
  a.cameras = cameras;
  a.points = points;
  a.results = results;
  a.records = records;
  a.work = static_cast<char*>(workspace);
  a.work_bytes = workspace_bytes;
  a.queue = static_cast<unsigned int*>(workspace);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_ba_batch_reset, dim3(1), dim3(64), 0, s, a.queue);
  CT_LAUNCH_CHECK();
  const int groups = batch->num_problems < cus ? batch->num_problems : cus;
  hipLaunchKernelGGL(k_ba_batch, dim3(groups), dim3(kBT), 0, s, a);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

// ---------------------------------------------------------------------------
// Host tables of a batch (the per-problem tables of ceres_mi355x.ba.build_tables,
// in one native pass per problem: counting sorts, O(N + T)).
// ---------------------------------------------------------------------------
namespace {

struct TableOut {
  int32_t* order;
  int32_t* cam;
  int32_t* pt;
  int32_t* point_offsets;
  int32_t* camera_offsets;
  int32_t* camera_obs;
  int32_t* block_offsets;
  int32_t* block_terms;
};

// One problem: observations grouped by (point, camera) with ties in input
// order, the CSR tables, and the observation pairs of each camera-pair block
// in ascending (a, b).  Returns the number of pairs; with out == nullptr it
// only counts.
int64_t problem_tables(int C, int P, int N, const int32_t* oc, const int32_t* op, const TableOut* out,
                       std::vector<int32_t>& scratch) {
  scratch.assign((size_t)3 * N + (size_t)P + (size_t)C + 2, 0);
  int32_t* by_cam = scratch.data();       // [N]
  int32_t* perm = by_cam + N;             // [N]
  int32_t* cams = perm + N;               // [N] camera of each sorted observation
  int32_t* cnt = cams + N;                // [max(P, C) + 2]
  // stable counting sort by camera, then by point: sorted by (point, camera), ties by input order
  std::fill(cnt, cnt + C + 1, 0);
  for (int o = 0; o < N; ++o) ++cnt[oc[o] + 1];
  for (int c = 0; c < C; ++c) cnt[c + 1] += cnt[c];
  for (int o = 0; o < N; ++o) by_cam[cnt[oc[o]]++] = o;
  std::vector<int32_t> poff((size_t)P + 1, 0);
  for (int o = 0; o < N; ++o) ++poff[op[o] + 1];
  for (int p = 0; p < P; ++p) poff[p + 1] += poff[p];
  {
    std::vector<int32_t> next(poff.begin(), poff.end() - 1);
    for (int k = 0; k < N; ++k) {
      const int o = by_cam[k];
      perm[next[op[o]]++] = o;
    }
  }
  for (int k = 0; k < N; ++k) cams[k] = oc[perm[k]];
  // block of pair (i, j), i <= j: row-major upper triangle
  auto block_of = [C](int i, int j) { return i * C - i * (i - 1) / 2 + (j - i); };
  const int B = C * (C + 1) / 2;
  std::vector<int64_t> bcount((size_t)B + 1, 0);
  int64_t T = 0;
  for (int p = 0; p < P; ++p)
    for (int a = poff[p]; a < poff[p + 1]; ++a)
      for (int b = poff[p]; b < poff[p + 1]; ++b)
        if (cams[a] <= cams[b]) {
          ++bcount[block_of(cams[a], cams[b]) + 1];
          ++T;
        }
  if (out == nullptr) return T;
  for (int k = 0; k < N; ++k) {
    out->order[k] = perm[k];
    out->cam[k] = cams[k];
    out->pt[k] = op[perm[k]];
  }
  for (int p = 0; p <= P; ++p) out->point_offsets[p] = poff[p];
  // camera CSR of the sorted observations (stable: ascending observation)
  std::vector<int32_t> coff((size_t)C + 1, 0);
  for (int k = 0; k < N; ++k) ++coff[cams[k] + 1];
  for (int c = 0; c < C; ++c) coff[c + 1] += coff[c];
  for (int c = 0; c <= C; ++c) out->camera_offsets[c] = coff[c];
  {
    std::vector<int32_t> next(coff.begin(), coff.end() - 1);
    for (int k = 0; k < N; ++k) out->camera_obs[next[cams[k]]++] = k;
  }
  for (int b = 0; b < B; ++b) bcount[b + 1] += bcount[b];
  for (int b = 0; b <= B; ++b) out->block_offsets[b] = (int32_t)bcount[b];
  std::vector<int64_t> next(bcount.begin(), bcount.end() - 1);
  for (int p = 0; p < P; ++p)
    for (int a = poff[p]; a < poff[p + 1]; ++a)
      for (int b = poff[p]; b < poff[p + 1]; ++b)
        if (cams[a] <= cams[b]) {
          const int64_t at = next[block_of(cams[a], cams[b])]++;
          out->block_terms[2 * at] = a;
          out->block_terms[2 * at + 1] = b;
        }
  return T;
}

}  // namespace

extern "C" int ct_ba_batch_tables(int32_t num_problems, const int32_t* num_cameras, const int32_t* num_points,
                                  const int32_t* num_observations, const int32_t* obs_camera,
                                  const int32_t* obs_point, int64_t* num_terms, int32_t* order,
                                  int32_t* sorted_obs_camera, int32_t* sorted_obs_point, int32_t* point_offsets,
                                  int32_t* camera_offsets, int32_t* camera_obs, int32_t* block_offsets,
                                  int32_t* block_terms) {
  clear_error();
  CT_CHECK_ARG(num_problems >= 0, "negative batch size");
  CT_CHECK_ARG(num_problems == 0 || (num_cameras && num_points && num_observations && num_terms),
               "null size array");
  const bool build = order != nullptr;
  if (build) {
    CT_CHECK_ARG(sorted_obs_camera && sorted_obs_point && point_offsets && camera_offsets && camera_obs &&
                     block_offsets && block_terms,
                 "null table array");
  }
  const int K = num_problems;
  // each problem's slice of every array
  std::vector<int64_t> ob(K + 1, 0), pb(K + 1, 0), cb(K + 1, 0), bb(K + 1, 0), tb(K + 1, 0);
  for (int k = 0; k < K; ++k) {
    const int C = num_cameras[k], P = num_points[k], N = num_observations[k];
    CT_CHECK_ARG(C >= 1 && C <= CT_BA_BATCH_MAX_CAMERAS, "a batched bundle adjustment has 1..21 cameras");
    CT_CHECK_ARG(P >= 0 && N >= 0, "negative problem size");
    ob[k + 1] = ob[k] + N;
    pb[k + 1] = pb[k] + P + 1;
    cb[k + 1] = cb[k] + C + 1;
    bb[k + 1] = bb[k] + C * (C + 1) / 2 + 1;
    if (build) {
      CT_CHECK_ARG(num_terms[k] >= 0, "negative num_terms");
      tb[k + 1] = tb[k] + num_terms[k];
    }
  }
  CT_CHECK_ARG(ob[K] == 0 || (obs_camera && obs_point), "null observation arrays");
  // problems are independent: worker threads take them from a shared counter
  std::atomic<int> next{0};
  std::atomic<int> bad{-1};  // first problem with an error
  std::string bad_msg;
  std::mutex bad_lock;
  auto work = [&]() {
    std::vector<int32_t> scratch;
    for (int k = next.fetch_add(1); k < K; k = next.fetch_add(1)) {
      const int C = num_cameras[k], P = num_points[k], N = num_observations[k];
      const int32_t* oc = obs_camera + ob[k];
      const int32_t* op = obs_point + ob[k];
      bool ok = true;
      for (int o = 0; o < N && ok; ++o) {
        if (oc[o] < 0 || oc[o] >= C || op[o] < 0 || op[o] >= P) {
          std::lock_guard<std::mutex> g(bad_lock);
          if (bad.load() < 0 || k < bad.load()) {
            bad.store(k);
            bad_msg = "problem " + std::to_string(k) + ": observation " + std::to_string(o) + " out of range";
          }
          ok = false;
        }
      }
      if (!ok) continue;
      if (build) {
        const TableOut out{order + ob[k],          sorted_obs_camera + ob[k], sorted_obs_point + ob[k],
                           point_offsets + pb[k],  camera_offsets + cb[k],    camera_obs + ob[k],
                           block_offsets + bb[k],  block_terms + 2 * tb[k]};
        if (problem_tables(C, P, N, oc, op, &out, scratch) != num_terms[k]) {
          std::lock_guard<std::mutex> g(bad_lock);
          if (bad.load() < 0 || k < bad.load()) {
            bad.store(k);
            bad_msg = "problem " + std::to_string(k) + ": num_terms differs from the counting call";
          }
        }
      } else {
        num_terms[k] = problem_tables(C, P, N, oc, op, nullptr, scratch);
      }
    }
  };
  const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const int nthreads = (int)std::min<int64_t>(hw, std::max(1, K));
  std::vector<std::thread> pool;
  for (int i = 1; i < nthreads; ++i) pool.emplace_back(work);
  work();
  for (std::thread& th : pool) th.join();
  if (bad.load() >= 0) {
    set_error(bad_msg);
    return CT_ERR_ARGUMENT;
  }
  return CT_OK;
}
