// SE2 pose-graph LM solve on the device (examples/pose_graph_2d.py:6-46; the
// trust-region loop of solver.py:35-357).
//
// Residual of an edge (a, b) with measurement m: relative_pose(a, b) - m,
// relative_pose = (R(theta_a)^T (p_b - p_a), wrap(theta_b - theta_a))
// (pose_graph_2d.py:10-15), with a robust loss per edge.  One node is held
// constant (pose_graph_2d.py:34).  The reference's DENSE_QR step is the
// solution of the damped normal equations, assembled deterministically from
// per-node diagonal blocks (incident edges in table order) and per-node-pair
// off-diagonal blocks (the pair's edges in table order), then solved in FP64
// by one of two exact methods:
//   dense:     the whole system, factored by the reduced-camera-system
//              Cholesky;
//   separator: the nodes touched by a non-chain edge (and the constant node)
//              form a separator; every other node lies in a run of
//              consecutive nodes joined only by chain edges (i, i + 1), whose
//              system is block tridiagonal and touches at most two separator
//              nodes.  Each run is factored on its own (one thread, block
//              Cholesky along the chain), its Schur complement onto its two
//              separator nodes is subtracted from the separator system, that
//              dense system is factored by the same Cholesky, and each run is
//              back-substituted.  The plan takes this form whenever some node
//              lies in a run; it factors 3 x (separator nodes) unknowns
//              densely instead of 3 x (nodes).
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "ba_kernels.h"
#include "lm_logic.h"
#include "lm_plan.h"

namespace ct {

namespace {

constexpr int kPgThreads = 256;

struct PGView {
  int32_t N, E, P, fixed, n;
  const int32_t* __restrict__ edges;
  const double* __restrict__ meas;
  const int32_t* __restrict__ node_offsets;
  const int32_t* __restrict__ node_edges;
  const int32_t* __restrict__ pair_nodes;
  const int32_t* __restrict__ pair_offsets;
  const int32_t* __restrict__ pair_edges;
  ct_loss loss;
};

struct PGWork {
  double* x;          // [2][N][3] ping-pong poses
  double* r;          // [E][3]    corrected residuals
  double* Ja;         // [E][9]    corrected d r / d pose a (row-major 3x3)
  double* Jb;         // [E][9]
  double* H;          // [n][n]    damped normal matrix
  double* rhs;        // [n]       -g
  double* dx;         // [n]
  double* part_cost;  // [Ge]
  double* part_grad;  // [Gn][2]   |g|^2, max|g|
  double* part_step;  // [Gn][3]   |dx|^2, |x'|^2, non-finite
  double* part_cand;  // [Ge][2]   candidate cost, model decrease
  double* records;
  LMState* state;
  double* solve_tmp;
};

// Separator form (see the file comment).  Blocks are 3x3 row-major.
struct PGSep {
  int32_t runs;   // runs of consecutive non-separator nodes
  int32_t nsep;   // separator nodes other than the constant one
  int32_t nent;   // off-diagonal separator blocks (joined pairs and runs with two ends)
  const int32_t* __restrict__ run_nodes;  // [runs][4] first node s, last node t, separator unknowns of the
                                          // left end L = s - 1 and right end R = t + 1 (-1: no coupling)
  const int32_t* __restrict__ run_pairs;  // [runs][2] pair of (L, s) and of (t, R), or -1
  const int32_t* __restrict__ chain_pair; // [N] for a run node k after its run's first: pair of (k - 1, k), or -1
  const int32_t* __restrict__ sep_nodes;  // [nsep] node of each separator unknown, increasing
  const int32_t* __restrict__ sep_runs;   // [nsep][2] run whose right end it is, run whose left end it is (-1: none)
  const int32_t* __restrict__ entries;    // [nent][4] separator unknowns u < v, pair (or -1), run (or -1)
  double* nodeblk;  // [N][9]   damped diagonal block of each free node
  double* pairblk;  // [P][9]   H(u, v) of each joined pair u < v (rows u, columns v)
  double* fac;      // [N][15]  per run node: L_kk^-1 (lower, packed 6) and L_k,k-1 (9)
  double* contrib;  // [runs][33] Schur terms onto the run's ends: LL (9), LR (9), RR (9), rhs L (3), rhs R (3)
  double* rhs_s;    // [3 nsep]
  double* x_s;      // [3 nsep]
};

__device__ __forceinline__ int col_of(const PGView& v, int k) { return 3 * (v.fixed >= 0 && k > v.fixed ? k - 1 : k); }

__device__ __forceinline__ void pose_residual(const double* a, const double* b, const double* m, double (&r)[3],
                                              double* ca_out, double* sa_out, double* dx_out, double* dy_out) {
  double sa, ca;
  sincos(a[2], &sa, &ca);
  const double dx = b[0] - a[0], dy = b[1] - a[1], dth = b[2] - a[2];
  double ss, cc;
  sincos(dth, &ss, &cc);
  r[0] = ca * dx + sa * dy - m[0];
  r[1] = -sa * dx + ca * dy - m[1];
  r[2] = atan2(ss, cc) - m[2];
  if (ca_out) {
    *ca_out = ca;
    *sa_out = sa;
    *dx_out = dx;
    *dy_out = dy;
  }
}

// Corrector of a 3-residual block (losses.py:158-193); returns 1/2 rho0.
__device__ __forceinline__ double robustify3(const ct_loss& L, double (&r)[3], double (&Ja)[9], double (&Jb)[9]) {
  const double s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
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
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double ta = alpha_s * (r[0] * Ja[j] + r[1] * Ja[3 + j] + r[2] * Ja[6 + j]);
    const double tb = alpha_s * (r[0] * Jb[j] + r[1] * Jb[3 + j] + r[2] * Jb[6 + j]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Ja[3 * k + j] = sq * (Ja[3 * k + j] - r[k] * ta);
      Jb[3 * k + j] = sq * (Jb[3 * k + j] - r[k] * tb);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) r[k] *= scale;
  return 0.5 * rho0;
}

__device__ __forceinline__ double robust_cost3(const ct_loss& L, const double (&r)[3]) {
  const double s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (L.type == CT_LOSS_TRIVIAL && L.scale == 1.0) return 0.5 * s;
  double rho1, rho2;
  return 0.5 * loss_rho(L, s, &rho1, &rho2);
}

__global__ __launch_bounds__(kPgThreads) void k_pg_init(PGView v, PGWork w, const double* __restrict__ poses,
                                                         ct_lm_options opt) {
  for (int i = blockIdx.x * kPgThreads + threadIdx.x; i < 3 * v.N; i += gridDim.x * kPgThreads) w.x[i] = poses[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) lm_state_init(w.state, opt);
}

__global__ __launch_bounds__(kPgThreads) void k_pg_finalize(PGView v, PGWork w, double* __restrict__ poses) {
  const int cur = w.state->cur;
  for (int i = blockIdx.x * kPgThreads + threadIdx.x; i < 3 * v.N; i += gridDim.x * kPgThreads)
    poses[i] = w.x[(size_t)cur * 3 * v.N + i];
}

// residuals and Jacobians of every edge at x; cost partials
__global__ __launch_bounds__(kPgThreads) void k_pg_eval(PGView v, PGWork w) {
  __shared__ double red[kPgThreads / kWave];
  const LMState* st = w.state;
  if (st->done) return;
  const double* x = w.x + (size_t)st->cur * 3 * v.N;
  const int e = blockIdx.x * kPgThreads + threadIdx.x;
  double cost = 0.0;
  if (e < v.E) {
    const int a = v.edges[2 * e], b = v.edges[2 * e + 1];
    double r[3], ca, sa, dx, dy;
    pose_residual(x + 3 * a, x + 3 * b, v.meas + 3 * e, r, &ca, &sa, &dx, &dy);
    double Ja[9] = {-ca, -sa, -sa * dx + ca * dy, sa, -ca, -ca * dx - sa * dy, 0.0, 0.0, -1.0};
    double Jb[9] = {ca, sa, 0.0, -sa, ca, 0.0, 0.0, 0.0, 1.0};
    cost = robustify3(v.loss, r, Ja, Jb);
#pragma unroll
    for (int k = 0; k < 3; ++k) w.r[3 * e + k] = r[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      w.Ja[9 * (size_t)e + k] = Ja[k];
      w.Jb[9 * (size_t)e + k] = Jb[k];
    }
  }
  const double t = block_sum<kPgThreads>(cost, red);
  if (threadIdx.x == 0) w.part_cost[blockIdx.x] = t;
}

__global__ __launch_bounds__(kPgThreads) void k_pg_zero(PGWork w, size_t count) {
  if (w.state->done) return;
  for (size_t i = blockIdx.x * (size_t)kPgThreads + threadIdx.x; i < count; i += (size_t)gridDim.x * kPgThreads)
    w.H[i] = 0.0;
}

// diagonal block, damping and gradient of each free node (incident edges in
// table order); into the dense H, or (separator form) into nodeblk
template <bool kSep>
__global__ __launch_bounds__(kPgThreads) void k_pg_node(PGView v, PGWork w, PGSep sp) {
  __shared__ double red[kPgThreads / kWave];
  const LMState* st = w.state;
  if (st->done) return;
  const int k = blockIdx.x * kPgThreads + threadIdx.x;
  double g2 = 0.0, gm = 0.0;
  if (k < v.N && k != v.fixed) {
    double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0};
    for (int t = v.node_offsets[k]; t < v.node_offsets[k + 1]; ++t) {
      const int tag = v.node_edges[t];
      const int e = tag >> 1;
      const double* J = ((tag & 1) ? w.Jb : w.Ja) + 9 * (size_t)e;
      const double* r = w.r + 3 * e;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) h[3 * i + j] += J[i] * J[j] + J[3 + i] * J[3 + j] + J[6 + i] * J[6 + j];
        g[i] += J[i] * r[0] + J[3 + i] * r[1] + J[6 + i] * r[2];
      }
    }
    const int c = col_of(v, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      // damping D = clamp(colsum J^2)/radius (linear.py:539-551)
      const double dg = st->jacobi_scaling ? h[4 * i] : 1.0;
      const double D = fmin(fmax(dg, st->min_diag), st->max_diag) / fmax(st->radius, 2.220446049250313e-16);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double hv = h[3 * i + j] + (i == j ? D : 0.0);
        if (kSep)
          sp.nodeblk[9 * (size_t)k + 3 * i + j] = hv;
        else
          w.H[(size_t)(c + i) * v.n + c + j] = hv;
      }
      w.rhs[c + i] = -g[i];
    }
    g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    gm = fmax(fabs(g[0]), fmax(fabs(g[1]), fabs(g[2])));
  }
  const double t = block_sum<kPgThreads>(g2, red);
  const double m = block_max<kPgThreads>(gm, red);
  if (threadIdx.x == 0) {
    w.part_grad[2 * blockIdx.x] = t;
    w.part_grad[2 * blockIdx.x + 1] = m;
  }
}

// off-diagonal block of each joined node pair u < v (its edges in table
// order); into the dense H, or (separator form) into pairblk
template <bool kSep>
__global__ __launch_bounds__(kPgThreads) void k_pg_pair(PGView v, PGWork w, PGSep sp) {
  if (w.state->done) return;
  const int p = blockIdx.x * kPgThreads + threadIdx.x;
  if (p >= v.P) return;
  const int u = v.pair_nodes[2 * p], q = v.pair_nodes[2 * p + 1];
  if (u == v.fixed || q == v.fixed) return;
  double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = v.pair_offsets[p]; t < v.pair_offsets[p + 1]; ++t) {
    const int tag = v.pair_edges[t];
    const int e = tag >> 1;
    const bool u_is_b = tag & 1;
    const double* Ju = (u_is_b ? w.Jb : w.Ja) + 9 * (size_t)e;
    const double* Jq = (u_is_b ? w.Ja : w.Jb) + 9 * (size_t)e;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) h[3 * i + j] += Ju[i] * Jq[j] + Ju[3 + i] * Jq[3 + j] + Ju[6 + i] * Jq[6 + j];
  }
  if (kSep) {
#pragma unroll
    for (int i = 0; i < 9; ++i) sp.pairblk[9 * (size_t)p + i] = h[i];
    return;
  }
  const int cu = col_of(v, u), cq = col_of(v, q);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      w.H[(size_t)(cu + i) * v.n + cq + j] = h[3 * i + j];
      w.H[(size_t)(cq + j) * v.n + cu + i] = h[3 * i + j];
    }
}

// ---------------------------------------------------------------------------
// Separator form
// ---------------------------------------------------------------------------

// 3x3 block Cholesky M = L L^T; Li = L^-1 (lower, packed l00 l10 l11 l20 l21
// l22).  False unless M is numerically positive definite.
__device__ __forceinline__ bool chol3_inverse(const double (&M)[9], double (&Li)[6]) {
  const double d0 = M[0];
  if (!(d0 > 0.0) || !isfinite(d0)) return false;
  const double l00 = sqrt(d0);
  const double l10 = M[3] / l00, l20 = M[6] / l00;
  const double d1 = M[4] - l10 * l10;
  if (!(d1 > 0.0) || !isfinite(d1)) return false;
  const double l11 = sqrt(d1);
  const double l21 = (M[7] - l20 * l10) / l11;
  const double d2 = M[8] - l20 * l20 - l21 * l21;
  if (!(d2 > 0.0) || !isfinite(d2)) return false;
  const double l22 = sqrt(d2);
  const double i00 = 1.0 / l00, i11 = 1.0 / l11, i22 = 1.0 / l22;
  const double i10 = -l10 * i00 * i11;
  const double i21 = -l21 * i11 * i22;
  const double i20 = -(l20 * i00 + l21 * i10) * i22;
  Li[0] = i00; Li[1] = i10; Li[2] = i11; Li[3] = i20; Li[4] = i21; Li[5] = i22;
  return true;
}

// y = Li x for the packed lower inverse
__device__ __forceinline__ void lower_apply(const double (&Li)[6], const double* x, double* y) {
  const double x0 = x[0], x1 = x[1], x2 = x[2];
  y[0] = Li[0] * x0;
  y[1] = Li[1] * x0 + Li[2] * x1;
  y[2] = Li[3] * x0 + Li[4] * x1 + Li[5] * x2;
}

// y = Li^T x
__device__ __forceinline__ void lower_apply_t(const double (&Li)[6], const double* x, double* y) {
  const double x0 = x[0], x1 = x[1], x2 = x[2];
  y[0] = Li[0] * x0 + Li[1] * x1 + Li[3] * x2;
  y[1] = Li[2] * x1 + Li[4] * x2;
  y[2] = Li[5] * x2;
}

// block of H(k, k - 1) (the transpose of the pair block H(k - 1, k)), or 0
__device__ __forceinline__ void block_t(const double* __restrict__ pairblk, int q, double (&B)[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) B[3 * i + j] = q >= 0 ? pairblk[9 * (size_t)q + 3 * j + i] : 0.0;
}

// One wavefront per run s..t: block Cholesky along the chain,
//   L_kk L_kk^T = A_k - L_k,k-1 L_k,k-1^T,  L_k,k-1 = H(k, k-1) L_k-1,k-1^-T,
// forward substitution of the couplings to the run's ends and of its
// right-hand side, W = L^-1 [H(run, L) | H(run, R) | b], and the Schur terms
// W^T W onto the ends.  The right coupling is non-zero only at t.  The
// recurrence is sequential along the chain: the wavefront stages kRunChunk
// nodes' blocks in LDS (one load per lane, all in flight together) and lane 0
// walks them.
constexpr int kRunChunk = 64;

__global__ __launch_bounds__(kRunChunk) void k_pg_runs(PGView v, PGWork w, PGSep sp) {
  __shared__ double sA[kRunChunk][9], sB[kRunChunk][9], sb[kRunChunk][3];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int r = blockIdx.x;  // the grid has exactly one workgroup per run
  const int lane = threadIdx.x;
  const int s = sp.run_nodes[4 * r], t = sp.run_nodes[4 * r + 1];
  const int L = sp.run_nodes[4 * r + 2], R = sp.run_nodes[4 * r + 3];
  const int pl = sp.run_pairs[2 * r], pr = sp.run_pairs[2 * r + 1];
  double WL[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, wb[3] = {0, 0, 0};  // this node's rows of W (left columns, b)
  double GLL[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, GLb[3] = {0, 0, 0};
  double Li[6] = {0, 0, 0, 0, 0, 0};
  bool ok = true;
  for (int c0 = s; c0 <= t; c0 += kRunChunk) {
    const int c1 = min(t, c0 + kRunChunk - 1);
    {
      const int k = c0 + lane;
      if (k <= c1) {
        double B[9];
        block_t(sp.pairblk, k > s ? sp.chain_pair[k] : -1, B);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          sA[lane][i] = sp.nodeblk[9 * (size_t)k + i];
          sB[lane][i] = B[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) sb[lane][i] = w.rhs[col_of(v, k) + i];
      }
    }
    __syncthreads();
    if (lane == 0) {
      for (int k = c0; k <= c1; ++k) {
        const int j = k - c0;
        double Ls[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (k > s) {
          // L_k,k-1 = B Li_prev^T: row i is Li_prev applied to row i of B
#pragma unroll
          for (int i = 0; i < 3; ++i) lower_apply(Li, &sB[j][3 * i], Ls + 3 * i);
        }
        double M[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj)
            M[3 * i + jj] = sA[j][3 * i + jj] -
                            (Ls[3 * i] * Ls[3 * jj] + Ls[3 * i + 1] * Ls[3 * jj + 1] + Ls[3 * i + 2] * Ls[3 * jj + 2]);
        if (!chol3_inverse(M, Li)) ok = false;
        double* f = sp.fac + 15 * (size_t)k;
#pragma unroll
        for (int i = 0; i < 6; ++i) f[i] = Li[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) f[6 + i] = Ls[i];
        // T = [H(k, L) | b_k] - L_k,k-1 [W_L | w_b]_(k-1), then W_k = Li T
        double TL[9], tb[3];
        double HL[9];
        block_t(sp.pairblk, k == s && L >= 0 ? pl : -1, HL);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
          for (int jj = 0; jj < 3; ++jj)
            TL[3 * i + jj] =
                HL[3 * i + jj] - (Ls[3 * i] * WL[jj] + Ls[3 * i + 1] * WL[3 + jj] + Ls[3 * i + 2] * WL[6 + jj]);
          tb[i] = sb[j][i] - (Ls[3 * i] * wb[0] + Ls[3 * i + 1] * wb[1] + Ls[3 * i + 2] * wb[2]);
        }
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
          const double col[3] = {TL[jj], TL[3 + jj], TL[6 + jj]};
          double out[3];
          lower_apply(Li, col, out);
          WL[jj] = out[0]; WL[3 + jj] = out[1]; WL[6 + jj] = out[2];
        }
        lower_apply(Li, tb, wb);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
          for (int jj = 0; jj < 3; ++jj)
            GLL[3 * i + jj] += WL[i] * WL[jj] + WL[3 + i] * WL[3 + jj] + WL[6 + i] * WL[6 + jj];
          GLb[i] += WL[i] * wb[0] + WL[3 + i] * wb[1] + WL[6 + i] * wb[2];
        }
      }
    }
    __syncthreads();
  }
  if (lane != 0) return;
  double* out = sp.contrib + 33 * (size_t)r;
  double WR[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (R >= 0) {
    // W_R at t = Li_t H(t, R); H(t, R) is the pair block itself (t < R)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double col[3] = {sp.pairblk[9 * (size_t)pr + j], sp.pairblk[9 * (size_t)pr + 3 + j],
                             sp.pairblk[9 * (size_t)pr + 6 + j]};
      double o3[3];
      lower_apply(Li, col, o3);
      WR[j] = o3[0]; WR[3 + j] = o3[1]; WR[6 + j] = o3[2];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      out[3 * i + j] = GLL[3 * i + j];
      out[9 + 3 * i + j] = WL[i] * WR[j] + WL[3 + i] * WR[3 + j] + WL[6 + i] * WR[6 + j];
      out[18 + 3 * i + j] = WR[i] * WR[j] + WR[3 + i] * WR[3 + j] + WR[6 + i] * WR[6 + j];
    }
    out[27 + i] = GLb[i];
    out[30 + i] = WR[i] * wb[0] + WR[3 + i] * wb[1] + WR[6 + i] * wb[2];
  }
  if (!ok) w.state->step_invalid = 1;
}

// The separator system: thread per separator unknown (its diagonal block and
// right-hand side: the node's block minus the Schur terms of the run ending
// at it and of the run starting at it, in that order), then thread per
// off-diagonal entry (the pair block minus the Schur term of the run between
// the two nodes), written with its transpose.  S holds zeros elsewhere.
__global__ __launch_bounds__(kPgThreads) void k_pg_sep_assemble(PGView v, PGWork w, PGSep sp) {
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int idx = blockIdx.x * kPgThreads + threadIdx.x;
  const int ns = 3 * sp.nsep;
  double* S = w.H;
  if (idx < sp.nsep) {
    const int k = sp.sep_nodes[idx];
    const int rb = sp.sep_runs[2 * idx], ra = sp.sep_runs[2 * idx + 1];
    const int c = col_of(v, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double h = sp.nodeblk[9 * (size_t)k + 3 * i + j];
        if (rb >= 0) h -= sp.contrib[33 * (size_t)rb + 18 + 3 * i + j];
        if (ra >= 0) h -= sp.contrib[33 * (size_t)ra + 3 * i + j];
        S[(size_t)(3 * idx + i) * ns + 3 * idx + j] = h;
      }
      double b = w.rhs[c + i];
      if (rb >= 0) b -= sp.contrib[33 * (size_t)rb + 30 + i];
      if (ra >= 0) b -= sp.contrib[33 * (size_t)ra + 27 + i];
      sp.rhs_s[3 * idx + i] = b;
    }
    return;
  }
  const int e = idx - sp.nsep;
  if (e >= sp.nent) return;
  const int u = sp.entries[4 * e], q = sp.entries[4 * e + 1], pq = sp.entries[4 * e + 2], rr = sp.entries[4 * e + 3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double h = pq >= 0 ? sp.pairblk[9 * (size_t)pq + 3 * i + j] : 0.0;
      if (rr >= 0) h -= sp.contrib[33 * (size_t)rr + 9 + 3 * i + j];
      S[(size_t)(3 * u + i) * ns + 3 * q + j] = h;
      S[(size_t)(3 * q + j) * ns + 3 * u + i] = h;
    }
}

// dx of every free node.  Workgroups [0, runs): one wavefront per run, as in
// k_pg_runs: forward substitution of b - H(run, ends) x_ends, then backward,
// each over LDS-staged chunks walked by lane 0 (the forward values pass
// through dx).  Workgroups from runs on copy the separator unknowns from x_s.
__global__ __launch_bounds__(kRunChunk) void k_pg_backsub(PGView v, PGWork w, PGSep sp) {
  __shared__ double sF[kRunChunk][15], sv[kRunChunk][3];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= sp.runs) {
    const int idx = ((int)blockIdx.x - sp.runs) * kRunChunk + lane;
    if (idx < sp.nsep) {
      const int c = col_of(v, sp.sep_nodes[idx]);
#pragma unroll
      for (int i = 0; i < 3; ++i) w.dx[c + i] = sp.x_s[3 * idx + i];
    }
    return;
  }
  const int r = blockIdx.x;
  const int s = sp.run_nodes[4 * r], t = sp.run_nodes[4 * r + 1];
  const int L = sp.run_nodes[4 * r + 2], R = sp.run_nodes[4 * r + 3];
  const int pl = sp.run_pairs[2 * r], pr = sp.run_pairs[2 * r + 1];
  // the ends' solution (from x_s: other workgroups of this launch write the
  // separator nodes' dx)
  double xl[3] = {0, 0, 0}, xr[3] = {0, 0, 0};
  if (L >= 0)
    for (int i = 0; i < 3; ++i) xl[i] = sp.x_s[3 * L + i];
  if (R >= 0)
    for (int i = 0; i < 3; ++i) xr[i] = sp.x_s[3 * R + i];
  // forward: v_k = Li_k (b_k - couplings - L_k,k-1 v_k-1)
  double vp[3] = {0, 0, 0};
  for (int c0 = s; c0 <= t; c0 += kRunChunk) {
    const int c1 = min(t, c0 + kRunChunk - 1);
    {
      const int k = c0 + lane;
      if (k <= c1) {
#pragma unroll
        for (int i = 0; i < 15; ++i) sF[lane][i] = sp.fac[15 * (size_t)k + i];
#pragma unroll
        for (int i = 0; i < 3; ++i) sv[lane][i] = w.rhs[col_of(v, k) + i];
      }
    }
    __syncthreads();
    if (lane == 0) {
      for (int k = c0; k <= c1; ++k) {
        const int j = k - c0;
        double Li[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) Li[i] = sF[j][i];
        const double* Ls = &sF[j][6];
        double tvec[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          tvec[i] = sv[j][i] - (Ls[3 * i] * vp[0] + Ls[3 * i + 1] * vp[1] + Ls[3 * i + 2] * vp[2]);
        if (k == s && L >= 0) {
          // - H(s, L) x_L, H(s, L) the transpose of the pair block H(L, s)
#pragma unroll
          for (int i = 0; i < 3; ++i)
            tvec[i] -= sp.pairblk[9 * (size_t)pl + i] * xl[0] + sp.pairblk[9 * (size_t)pl + 3 + i] * xl[1] +
                       sp.pairblk[9 * (size_t)pl + 6 + i] * xl[2];
        }
        if (k == t && R >= 0) {
#pragma unroll
          for (int i = 0; i < 3; ++i)
            tvec[i] -= sp.pairblk[9 * (size_t)pr + 3 * i] * xr[0] + sp.pairblk[9 * (size_t)pr + 3 * i + 1] * xr[1] +
                       sp.pairblk[9 * (size_t)pr + 3 * i + 2] * xr[2];
        }
        lower_apply(Li, tvec, vp);
        const int c = col_of(v, k);
#pragma unroll
        for (int i = 0; i < 3; ++i) w.dx[c + i] = vp[i];
      }
    }
    __syncthreads();
  }
  // backward, chunks from the end: x_k = Li_k^T (v_k - L_k+1,k^T x_k+1)
  double xn[3] = {0, 0, 0};
  double Lsn[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // L_k+1,k
  for (int e1 = t; e1 >= s; e1 -= kRunChunk) {
    const int e0 = max(s, e1 - kRunChunk + 1);
    {
      const int k = e0 + lane;
      if (k <= e1) {
#pragma unroll
        for (int i = 0; i < 15; ++i) sF[lane][i] = sp.fac[15 * (size_t)k + i];
#pragma unroll
        for (int i = 0; i < 3; ++i) sv[lane][i] = w.dx[col_of(v, k) + i];  // v_k, written by lane 0 above
      }
    }
    __syncthreads();
    if (lane == 0) {
      for (int k = e1; k >= e0; --k) {
        const int j = k - e0;
        double Li[6], tvec[3];
#pragma unroll
        for (int i = 0; i < 6; ++i) Li[i] = sF[j][i];
#pragma unroll
        for (int i = 0; i < 3; ++i) tvec[i] = sv[j][i];
        if (k < t) {
#pragma unroll
          for (int i = 0; i < 3; ++i) tvec[i] -= Lsn[i] * xn[0] + Lsn[3 + i] * xn[1] + Lsn[6 + i] * xn[2];
        }
        lower_apply_t(Li, tvec, xn);
        const int c = col_of(v, k);
#pragma unroll
        for (int i = 0; i < 3; ++i) w.dx[c + i] = xn[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) Lsn[i] = sF[j][6 + i];
      }
    }
    __syncthreads();
  }
}

struct PGCounts {
  int ge, gn;  // partial rows of edge and node kernels
};

__global__ __launch_bounds__(kPgThreads) void k_pg_begin(PGCounts c, PGWork w) {
  __shared__ double red[kPgThreads / kWave];
  LMState* st = w.state;
  if (st->done) return;
  double cs = 0.0, g2 = 0.0, gm = 0.0;
  for (int i = threadIdx.x; i < c.ge; i += kPgThreads) cs += w.part_cost[i];
  for (int i = threadIdx.x; i < c.gn; i += kPgThreads) {
    g2 += w.part_grad[2 * i];
    gm = fmax(gm, w.part_grad[2 * i + 1]);
  }
  const double cost = block_sum<kPgThreads>(cs, red);
  const double gnorm2 = block_sum<kPgThreads>(g2, red);
  const double gmax = block_max<kPgThreads>(gm, red);
  if (threadIdx.x == 0) lm_begin_update(st, cost, gnorm2, gmax, w.records);
}

// x' = x + dx on the free nodes; |dx|^2, |x'|^2 over all nodes, non-finite count
__global__ __launch_bounds__(kPgThreads) void k_pg_step(PGView v, PGWork w) {
  __shared__ double red[kPgThreads / kWave];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const int cur = st->cur;
  const double* x = w.x + (size_t)cur * 3 * v.N;
  double* xn = w.x + (size_t)(1 - cur) * 3 * v.N;
  const int k = blockIdx.x * kPgThreads + threadIdx.x;
  double dx2 = 0.0, x2 = 0.0, bad = 0.0;
  if (k < v.N) {
    const int c = col_of(v, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double d = (k == v.fixed) ? 0.0 : w.dx[c + i];
      const double nv = x[3 * k + i] + d;
      xn[3 * k + i] = nv;
      dx2 += d * d;
      x2 += nv * nv;
      if (!isfinite(d)) bad = 1.0;
    }
  }
  const double a0 = block_sum<kPgThreads>(dx2, red);
  const double a1 = block_sum<kPgThreads>(x2, red);
  const double a2 = block_sum<kPgThreads>(bad, red);
  if (threadIdx.x == 0) {
    w.part_step[3 * blockIdx.x] = a0;
    w.part_step[3 * blockIdx.x + 1] = a1;
    w.part_step[3 * blockIdx.x + 2] = a2;
  }
}

// candidate cost at x' and the model decrease -r.u - u.u/2, u = Ja dx_a + Jb dx_b
__global__ __launch_bounds__(kPgThreads) void k_pg_cand(PGView v, PGWork w) {
  __shared__ double red[kPgThreads / kWave];
  const LMState* st = w.state;
  if (st->done || st->skip_step) return;
  const double* xn = w.x + (size_t)(1 - st->cur) * 3 * v.N;
  const int e = blockIdx.x * kPgThreads + threadIdx.x;
  double cost = 0.0, md = 0.0;
  if (e < v.E) {
    const int a = v.edges[2 * e], b = v.edges[2 * e + 1];
    double da[3] = {0.0, 0.0, 0.0}, db[3] = {0.0, 0.0, 0.0};
    if (a != v.fixed) {
      const int c = col_of(v, a);
      da[0] = w.dx[c];
      da[1] = w.dx[c + 1];
      da[2] = w.dx[c + 2];
    }
    if (b != v.fixed) {
      const int c = col_of(v, b);
      db[0] = w.dx[c];
      db[1] = w.dx[c + 1];
      db[2] = w.dx[c + 2];
    }
    const double* Ja = w.Ja + 9 * (size_t)e;
    const double* Jb = w.Jb + 9 * (size_t)e;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double u = Ja[3 * k] * da[0] + Ja[3 * k + 1] * da[1] + Ja[3 * k + 2] * da[2] + Jb[3 * k] * db[0] +
                       Jb[3 * k + 1] * db[1] + Jb[3 * k + 2] * db[2];
      md += -w.r[3 * e + k] * u - 0.5 * u * u;
    }
    double r[3];
    pose_residual(xn + 3 * a, xn + 3 * b, v.meas + 3 * e, r, nullptr, nullptr, nullptr, nullptr);
    cost = robust_cost3(v.loss, r);
  }
  const double a0 = block_sum<kPgThreads>(cost, red);
  const double a1 = block_sum<kPgThreads>(md, red);
  if (threadIdx.x == 0) {
    w.part_cand[2 * blockIdx.x] = a0;
    w.part_cand[2 * blockIdx.x + 1] = a1;
  }
}

__global__ __launch_bounds__(kPgThreads) void k_pg_end(PGCounts c, PGWork w) {
  __shared__ double red[kPgThreads / kWave];
  LMState* st = w.state;
  const int done = st->done;
  const int skip = st->skip_step;
  __syncthreads();
  if (done) return;
  if (skip) {
    if (threadIdx.x == 0) lm_skip_update(st);
    return;
  }
  double a[5] = {0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < c.ge; i += kPgThreads) {
    a[0] += w.part_cand[2 * i];
    a[1] += w.part_cand[2 * i + 1];
  }
  for (int i = threadIdx.x; i < c.gn; i += kPgThreads) {
    a[2] += w.part_step[3 * i];
    a[3] += w.part_step[3 * i + 1];
    a[4] += w.part_step[3 * i + 2];
  }
  double s[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k] = block_sum<kPgThreads>(a[k], red);
  if (threadIdx.x == 0) lm_end_update(st, s[0], s[1], s[2], s[3], s[4] != 0.0, w.records);
}

int blocks_of(int n) { return n > 0 ? (n + kPgThreads - 1) / kPgThreads : 1; }

struct PGLayout {
  PGWork w{};
  PGSep sep{};
  // separator tables, uploaded at plan creation (sizes bounded by N and P)
  int32_t* run_nodes = nullptr;   // [N][4]
  int32_t* run_pairs = nullptr;   // [N][2]
  int32_t* chain_pair = nullptr;  // [N]
  int32_t* sep_nodes = nullptr;   // [N]
  int32_t* sep_runs = nullptr;    // [N][2]
  int32_t* entries = nullptr;     // [P + N][4]
  size_t bytes = 0;
};

PGLayout pg_carve(const ct_pose2d_problem* p, const ct_lm_options* o, void* base) {
  PGLayout L;
  Carver cv(base);
  const size_t N = p->num_nodes, E = p->num_edges;
  const int n = 3 * (p->num_nodes - (p->fixed_node >= 0 ? 1 : 0));
  L.w.x = cv.take<double>(2 * 3 * N);
  L.w.r = cv.take<double>(3 * E);
  L.w.Ja = cv.take<double>(9 * E);
  L.w.Jb = cv.take<double>(9 * E);
  L.w.H = cv.take<double>((size_t)n * n);
  L.w.rhs = cv.take<double>((size_t)n);
  L.w.dx = cv.take<double>((size_t)n);
  L.w.part_cost = cv.take<double>((size_t)blocks_of((int)E) + 1);
  L.w.part_grad = cv.take<double>(2 * (size_t)blocks_of((int)N) + 2);
  L.w.part_step = cv.take<double>(3 * (size_t)blocks_of((int)N) + 3);
  L.w.part_cand = cv.take<double>(2 * (size_t)blocks_of((int)E) + 2);
  L.w.records = cv.take<double>((size_t)(o->max_num_iterations + 2) * CT_REC_FIELDS);
  L.w.state = cv.take<LMState>(1);
  L.w.solve_tmp = cv.take<double>(rcs_cholesky_tmp_doubles(n) + 1);
  // separator form: the separator system lives in H (3 nsep <= n)
  const size_t P = p->num_pairs;
  L.sep.nodeblk = cv.take<double>(9 * N);
  L.sep.pairblk = cv.take<double>(9 * P + 1);
  L.sep.fac = cv.take<double>(15 * N);
  L.sep.contrib = cv.take<double>(33 * N);
  L.sep.rhs_s = cv.take<double>(3 * N);
  L.sep.x_s = cv.take<double>(3 * N);
  L.run_nodes = cv.take<int32_t>(4 * N);
  L.run_pairs = cv.take<int32_t>(2 * N);
  L.chain_pair = cv.take<int32_t>(N);
  L.sep_nodes = cv.take<int32_t>(N);
  L.sep_runs = cv.take<int32_t>(2 * N);
  L.entries = cv.take<int32_t>(4 * (P + N));
  L.bytes = cv.offset + 256;
  return L;
}

// Separator form of the plan (see the file comment): which nodes are
// separator nodes, the runs between them, and the off-diagonal separator
// blocks.  *use is false when no node lies in a run (the dense form then).
int build_separator(const ct_pose2d_problem& p, const std::vector<int32_t>& ed, const std::vector<int32_t>& pn,
                    PGLayout& L, bool* use, int* nsep_out, hipStream_t s) {
  *use = false;
  const int N = p.num_nodes, E = p.num_edges, P = p.num_pairs, fixed = p.fixed_node;
  std::vector<char> is_sep(N, 0);
  if (fixed >= 0) is_sep[fixed] = 1;
  for (int e = 0; e < E; ++e) {
    const int a = ed[2 * e], b = ed[2 * e + 1];
    if (a - b != 1 && b - a != 1) is_sep[a] = is_sep[b] = 1;
  }
  int run_nodes_total = 0;
  for (int k = 0; k < N; ++k) run_nodes_total += !is_sep[k];
  if (run_nodes_total == 0) return CT_OK;
  // pair index of each joined pair u < v
  std::unordered_map<int64_t, int> pair_of;
  pair_of.reserve(2 * (size_t)P + 1);
  for (int q = 0; q < P; ++q) pair_of[(int64_t)pn[2 * q] * N + pn[2 * q + 1]] = q;
  auto find_pair = [&](int u, int v) {
    const auto it = pair_of.find((int64_t)u * N + v);
    return it == pair_of.end() ? -1 : it->second;
  };
  std::vector<int32_t> sep_index(N, -1), sep_nodes;
  for (int k = 0; k < N; ++k)
    if (is_sep[k] && k != fixed) {
      sep_index[k] = (int32_t)sep_nodes.size();
      sep_nodes.push_back(k);
    }
  const int nsep = (int)sep_nodes.size();
  std::vector<int32_t> run_nodes, run_pairs, chain_pair(N, -1), sep_runs(2 * (size_t)nsep, -1);
  // (u, v) separator unknowns -> (pair, run)
  std::map<std::pair<int, int>, std::pair<int, int>> ent;
  for (int q = 0; q < P; ++q) {
    const int u = sep_index[pn[2 * q]], v = sep_index[pn[2 * q + 1]];
    if (u >= 0 && v >= 0) ent[{u, v}] = {q, -1};
  }
  for (int k = 0; k < N;) {
    if (is_sep[k]) {
      ++k;
      continue;
    }
    const int s = k;
    while (k + 1 < N && !is_sep[k + 1]) ++k;
    const int t = k++;
    const int r = (int)(run_nodes.size() / 4);
    for (int j = s + 1; j <= t; ++j) chain_pair[j] = find_pair(j - 1, j);
    // an end couples when it is a separator unknown joined to the run by an edge
    int pl = s > 0 && sep_index[s - 1] >= 0 ? find_pair(s - 1, s) : -1;
    int pr = t + 1 < N && sep_index[t + 1] >= 0 ? find_pair(t, t + 1) : -1;
    const int sl = pl >= 0 ? sep_index[s - 1] : -1, sr = pr >= 0 ? sep_index[t + 1] : -1;
    run_nodes.insert(run_nodes.end(), {s, t, sl, sr});
    run_pairs.insert(run_pairs.end(), {pl, pr});
    if (sl >= 0) sep_runs[2 * (size_t)sl + 1] = r;
    if (sr >= 0) sep_runs[2 * (size_t)sr] = r;
    if (sl >= 0 && sr >= 0) {
      auto it = ent.find({sl, sr});
      if (it == ent.end())
        ent[{sl, sr}] = {-1, r};
      else
        it->second.second = r;
    }
  }
  std::vector<int32_t> entries;
  for (const auto& kv : ent) entries.insert(entries.end(), {kv.first.first, kv.first.second, kv.second.first, kv.second.second});
  const std::pair<int32_t*, const std::vector<int32_t>*> uploads[] = {
      {L.run_nodes, &run_nodes}, {L.run_pairs, &run_pairs}, {L.chain_pair, &chain_pair},
      {L.sep_nodes, &sep_nodes}, {L.sep_runs, &sep_runs},   {L.entries, &entries}};
  // stream-ordered copies (a legacy-stream copy would wait on, and disturb a graph capture in, other streams)
  for (const auto& u : uploads)
    if (!u.second->empty())
      CT_HIP(hipMemcpyAsync(u.first, u.second->data(), sizeof(int32_t) * u.second->size(), hipMemcpyHostToDevice, s));
  CT_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
  PGSep& sp = L.sep;
  sp.runs = (int32_t)(run_nodes.size() / 4);
  sp.nsep = nsep;
  sp.nent = (int32_t)(entries.size() / 4);
  sp.run_nodes = L.run_nodes;
  sp.run_pairs = L.run_pairs;
  sp.chain_pair = L.chain_pair;
  sp.sep_nodes = L.sep_nodes;
  sp.sep_runs = L.sep_runs;
  sp.entries = L.entries;
  *nsep_out = nsep;
  *use = true;
  return CT_OK;
}

int validate_pg(const ct_pose2d_problem* p) {
  CT_CHECK_ARG(p != nullptr, "problem is null");
  CT_CHECK_ARG(p->num_nodes >= 1 && p->num_edges >= 0 && p->num_pairs >= 0, "bad problem size");
  CT_CHECK_ARG(p->fixed_node >= -1 && p->fixed_node < p->num_nodes, "fixed_node out of range");
  CT_CHECK_ARG((int64_t)p->num_nodes * 3 <= 24576, "too many nodes for the dense normal equations");
  CT_CHECK_ARG(p->loss.type >= CT_LOSS_TRIVIAL && p->loss.type <= CT_LOSS_TUKEY, "unknown loss type");
  if (p->num_edges > 0) {
    CT_CHECK_ARG(p->edges && p->measurements && p->node_offsets && p->node_edges, "edge arrays are null");
    CT_CHECK_ARG(p->num_pairs == 0 || (p->pair_nodes && p->pair_offsets && p->pair_edges), "pair arrays are null");
  }
  return CT_OK;
}

template <typename T>
int copy_back(std::vector<T>& h, const T* d, size_t count, hipStream_t s) {
  h.resize(count);
  if (count) CT_HIP(hipMemcpyAsync(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost, s));
  return CT_OK;
}

// every index the kernels follow, checked once on the host; ed and pn keep
// the host copies of the edges and pair nodes
int validate_pg_tables(const ct_pose2d_problem& p, std::vector<int32_t>& ed, std::vector<int32_t>& pn, hipStream_t s) {
  const int N = p.num_nodes, E = p.num_edges, P = p.num_pairs;
  if (E == 0) return CT_OK;
  std::vector<int32_t> no, ne, po, pe;
  CT_TRY(copy_back(ed, p.edges, 2 * (size_t)E, s));
  CT_TRY(copy_back(no, p.node_offsets, (size_t)N + 1, s));
  CT_TRY(copy_back(ne, p.node_edges, 2 * (size_t)E, s));
  CT_TRY(copy_back(pn, p.pair_nodes, 2 * (size_t)P, s));
  CT_TRY(copy_back(po, p.pair_offsets, (size_t)P + 1, s));
  CT_TRY(copy_back(pe, p.pair_edges, (size_t)E, s));
  CT_HIP(hipStreamSynchronize(s));
  for (int e = 0; e < E; ++e)
    CT_CHECK_ARG(ed[2 * e] >= 0 && ed[2 * e] < N && ed[2 * e + 1] >= 0 && ed[2 * e + 1] < N &&
                     ed[2 * e] != ed[2 * e + 1],
                 "edges must join two distinct valid nodes");
  CT_CHECK_ARG(no[0] == 0 && no[N] == 2 * E, "node_offsets must cover 2 num_edges entries");
  for (int k = 0; k < N; ++k) CT_CHECK_ARG(no[k] <= no[k + 1], "node_offsets must be non-decreasing");
  for (int k = 0; k < N; ++k)
    for (int t = no[k]; t < no[k + 1]; ++t) {
      const int e = ne[t] >> 1;
      CT_CHECK_ARG(ne[t] >= 0 && e < E && ed[2 * e + (ne[t] & 1)] == k, "node_edges must list incident edges");
    }
  CT_CHECK_ARG(P == 0 || (po[0] == 0 && po[P] == E), "pair_offsets must cover num_edges entries");
  for (int q = 0; q < P; ++q) {
    CT_CHECK_ARG(po[q] <= po[q + 1], "pair_offsets must be non-decreasing");
    const int u = pn[2 * q], w = pn[2 * q + 1];
    CT_CHECK_ARG(u >= 0 && w < N && u < w, "pair_nodes must hold u < v");
    for (int t = po[q]; t < po[q + 1]; ++t) {
      const int e = pe[t] >> 1;
      CT_CHECK_ARG(pe[t] >= 0 && e < E, "pair_edges out of range");
      const bool u_is_b = pe[t] & 1;
      CT_CHECK_ARG(ed[2 * e + (u_is_b ? 1 : 0)] == u && ed[2 * e + (u_is_b ? 0 : 1)] == w,
                   "pair_edges must join the pair's nodes");
    }
  }
  return CT_OK;
}

}  // namespace

}  // namespace ct

using namespace ct;

struct ct_pose2d_plan {
  ct_pose2d_problem problem;
  ct_lm_options options;
  PGView view;
  PGLayout layout;
  PGCounts counts;
  bool separator = false;  // the separator form (else the dense form)
  int solve_n = 0;         // unknowns of the system factored densely per iteration
  double* user_poses = nullptr;
  hipStream_t stream = nullptr;
  LMDriver driver;
};

namespace {

int pg_enqueue(ct_pose2d_plan* pl, hipStream_t s, bool evaluation_only) {
  const PGView& v = pl->view;
  PGWork& w = pl->layout.w;
  const PGSep& sp = pl->layout.sep;
  const bool sep = pl->separator;
  if (v.E > 0) {
    hipLaunchKernelGGL(k_pg_eval, dim3(blocks_of(v.E)), dim3(kPgThreads), 0, s, v, w);
    CT_LAUNCH_CHECK();
  }
  // the dense system, or the separator system, starts from zeros
  if (pl->solve_n > 0) {
    const size_t count = (size_t)pl->solve_n * pl->solve_n;
    int grid = (int)((count + kPgThreads - 1) / kPgThreads);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_pg_zero, dim3(grid), dim3(kPgThreads), 0, s, w, count);
    CT_LAUNCH_CHECK();
  }
  if (sep)
    hipLaunchKernelGGL(k_pg_node<true>, dim3(blocks_of(v.N)), dim3(kPgThreads), 0, s, v, w, sp);
  else
    hipLaunchKernelGGL(k_pg_node<false>, dim3(blocks_of(v.N)), dim3(kPgThreads), 0, s, v, w, sp);
  CT_LAUNCH_CHECK();
  if (v.P > 0) {
    if (sep)
      hipLaunchKernelGGL(k_pg_pair<true>, dim3(blocks_of(v.P)), dim3(kPgThreads), 0, s, v, w, sp);
    else
      hipLaunchKernelGGL(k_pg_pair<false>, dim3(blocks_of(v.P)), dim3(kPgThreads), 0, s, v, w, sp);
    CT_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_pg_begin, dim3(1), dim3(kPgThreads), 0, s, pl->counts, w);
  CT_LAUNCH_CHECK();
  if (!evaluation_only && sep) {
    if (sp.runs > 0) {
      hipLaunchKernelGGL(k_pg_runs, dim3(sp.runs), dim3(kRunChunk), 0, s, v, w, sp);
      CT_LAUNCH_CHECK();
    }
    if (sp.nsep > 0) {
      hipLaunchKernelGGL(k_pg_sep_assemble, dim3(blocks_of(sp.nsep + sp.nent)), dim3(kPgThreads), 0, s, v, w, sp);
      CT_LAUNCH_CHECK();
      CT_TRY(launch_rcs_cholesky_solve(w.H, sp.rhs_s, sp.x_s, 3 * sp.nsep, w.state, w.solve_tmp, s,
                                        pl->options.solver_workgroups));
    }
    const int sep_groups = (sp.nsep + kRunChunk - 1) / kRunChunk;
    if (sp.runs + sep_groups > 0)
      hipLaunchKernelGGL(k_pg_backsub, dim3(sp.runs + sep_groups), dim3(kRunChunk), 0, s, v, w, sp);
    CT_LAUNCH_CHECK();
  }
  if (!evaluation_only) {
    if (!sep && v.n > 0)
      CT_TRY(launch_rcs_cholesky_solve(w.H, w.rhs, w.dx, v.n, w.state, w.solve_tmp, s, pl->options.solver_workgroups));
    hipLaunchKernelGGL(k_pg_step, dim3(blocks_of(v.N)), dim3(kPgThreads), 0, s, v, w);
    CT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pg_cand, dim3(blocks_of(v.E)), dim3(kPgThreads), 0, s, v, w);
    CT_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_pg_end, dim3(1), dim3(kPgThreads), 0, s, pl->counts, w);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace

extern "C" {

int ct_pose2d_workspace_bytes(const ct_pose2d_problem* problem, const ct_lm_options* options, size_t* bytes) {
  clear_error();
  // sizes only: a caller may ask before its tables exist on the device
  CT_CHECK_ARG(problem != nullptr, "problem is null");
  CT_CHECK_ARG(problem->num_nodes >= 1 && problem->num_edges >= 0 && problem->num_pairs >= 0, "bad problem size");
  CT_CHECK_ARG((int64_t)problem->num_nodes * 3 <= 24576, "too many nodes for the dense normal equations");
  CT_CHECK_ARG(options != nullptr && bytes != nullptr, "null argument");
  *bytes = pg_carve(problem, options, nullptr).bytes;
  return CT_OK;
}

int ct_pose2d_plan_create(const ct_pose2d_problem* problem, const ct_lm_options* options, double* poses,
                          void* workspace, size_t workspace_bytes, ct_stream_t stream, ct_pose2d_plan** plan) {
  clear_error();
  CT_TRY(validate_pg(problem));
  CT_CHECK_ARG(options != nullptr && plan != nullptr && poses != nullptr && workspace != nullptr, "null argument");
  CT_CHECK_ARG(options->max_num_iterations >= 0 && options->initial_trust_region_radius > 0.0, "bad options");
  CT_CHECK_ARG(options->loop_mode == CT_LOOP_HOST || options->loop_mode == CT_LOOP_DEVICE, "unknown loop mode");
  if (options->trust_region_strategy != CT_LEVENBERG_MARQUARDT) {
    ::ct::set_error("DOGLEG is on the device path for bundle adjustment only");
    return CT_ERR_UNSUPPORTED;
  }
  if (options->debug_invalid_steps != 0) {
    ::ct::set_error("debug_invalid_steps is a bundle-adjustment test hook");
    return CT_ERR_UNSUPPORTED;
  }
  const size_t need = pg_carve(problem, options, nullptr).bytes;
  if (workspace_bytes < need) {
    set_error("workspace too small: need " + std::to_string(need) + " bytes");
    return CT_ERR_WORKSPACE;
  }
  // the tables are read back on the caller's stream, behind its uploads
  std::vector<int32_t> ed, pn;
  CT_TRY(validate_pg_tables(*problem, ed, pn, reinterpret_cast<hipStream_t>(stream)));
  auto pl = std::make_unique<ct_pose2d_plan>();
  pl->problem = *problem;
  pl->options = *options;
  pl->layout = pg_carve(problem, options, workspace);
  PGView& v = pl->view;
  v.N = problem->num_nodes;
  v.E = problem->num_edges;
  v.P = problem->num_pairs;
  v.fixed = problem->fixed_node;
  v.n = 3 * (v.N - (v.fixed >= 0 ? 1 : 0));
  v.edges = problem->edges;
  v.meas = problem->measurements;
  v.node_offsets = problem->node_offsets;
  v.node_edges = problem->node_edges;
  v.pair_nodes = problem->pair_nodes;
  v.pair_offsets = problem->pair_offsets;
  v.pair_edges = problem->pair_edges;
  v.loss = problem->loss;
  pl->counts.ge = blocks_of(v.E);
  pl->counts.gn = blocks_of(v.N);
  pl->user_poses = poses;
  pl->stream = reinterpret_cast<hipStream_t>(stream);
  int nsep = 0;
  CT_TRY(build_separator(*problem, ed, pn, pl->layout, &pl->separator, &nsep, pl->stream));
  pl->solve_n = pl->separator ? 3 * nsep : v.n;
  if (pl->solve_n > 0) {
    CT_TRY(rcs_cholesky_prepare(pl->layout.w.solve_tmp, pl->solve_n, pl->stream));
    CT_HIP(hipStreamSynchronize(pl->stream));
  }
  ct_pose2d_plan* raw = pl.get();
  LMDriver& d = pl->driver;
  d.enqueue_iteration = [raw](hipStream_t s, bool evaluation_only) { return pg_enqueue(raw, s, evaluation_only); };
  d.launch_init = [raw](hipStream_t s) {
    hipLaunchKernelGGL(k_pg_init, dim3(blocks_of(3 * raw->view.N)), dim3(kPgThreads), 0, s, raw->view,
                       raw->layout.w, raw->user_poses, raw->options);
    CT_LAUNCH_CHECK();
    return CT_OK;
  };
  d.launch_finalize = [raw](hipStream_t s) {
    hipLaunchKernelGGL(k_pg_finalize, dim3(blocks_of(3 * raw->view.N)), dim3(kPgThreads), 0, s, raw->view,
                       raw->layout.w, raw->user_poses);
    CT_LAUNCH_CHECK();
    return CT_OK;
  };
  const int st = lm_driver_setup(d, *options, pl->stream, pl->layout.w.state, pl->layout.w.records);
  if (st != CT_OK) {
    lm_driver_release(d);
    return st;
  }
  *plan = pl.release();
  return CT_OK;
}

int ct_pose2d_plan_run(ct_pose2d_plan* plan, ct_lm_report* report, double* records) {
  clear_error();
  CT_CHECK_ARG(plan != nullptr, "plan is null");
  return lm_driver_run(plan->driver, report, records);
}

int ct_pose2d_plan_system_size(const ct_pose2d_plan* plan, int32_t* unknowns) {
  clear_error();
  CT_CHECK_ARG(plan != nullptr && unknowns != nullptr, "null argument");
  *unknowns = plan->solve_n;
  return CT_OK;
}

int ct_pose2d_plan_destroy(ct_pose2d_plan* plan) {
  clear_error();
  if (!plan) return CT_OK;
  lm_driver_release(plan->driver);
  delete plan;
  return CT_OK;
}

}  // extern "C"
