// Covariance of a bundle-adjustment solution from the factor of the reduced
// camera system (the computation behind covariance.py:86-172 for a
// full-rank Jacobian, where the reference's (J^T J)^+ is (J^T J)^-1).
//
// With H = J^T J = [[V, W], [W^T, U]] (points first, V block diagonal) and the
// Schur complement S = U - W^T V^-1 W:
//   (H^-1)_cc = S^-1 = L^-T L^-1            (S = L L^T, the FP64 Cholesky)
//   (H^-1)_pc = -E_p S^-1,                  E_p = V_p^-1 W_p
//   (H^-1)_pq = [p = q] V_p^-1 + E_p S^-1 E_q^T
// E_p has one 3x9 block per observation o of point p, at the observing
// camera: F_o = V_p^-1 Jp_o^T Jc_o = L_p^-T P_o^T Jc_o, with L_p the point's
// 3x3 factor and P_o = Jp_o L_p^-T, both already in the evaluation's buffers.
// So only S^-1 (dense, 9C x 9C) is formed; each requested block is then a
// short sum over the observations of its points.
#include "ba_kernels.h"

namespace ct {

namespace {

constexpr int kCovTile = 64;  // X = Y^T Y tile
constexpr int kCovK = 16;     // rows of Y per LDS stage

// Y = L^-1 for the lower factor L (row-major n x n): one wavefront per column
// j solves L y = e_j by forward substitution, row by row (each row's dot
// product reads L's row contiguously).  Only entries i >= j are written.
__global__ __launch_bounds__(kWave) void k_cov_lower_inverse(const double* __restrict__ L, int n,
                                                             double* __restrict__ Y, const LMState* st) {
  extern __shared__ double y[];
  if (st->step_invalid) return;  // the factorisation failed: the status reports it
  const int j = blockIdx.x;
  const int lane = threadIdx.x;
  for (int i = j; i < n; ++i) {
    const double* row = L + (size_t)i * n;
    double part = 0.0;
    for (int k = j + lane; k < i; k += kWave) part = fma(row[k], y[k], part);
    const double dot = wave_sum(part);
    const double yi = ((i == j ? 1.0 : 0.0) - dot) / row[i];
    if (lane == 0) {
      y[i] = yi;
      Y[(size_t)i * n + j] = yi;
    }
    __syncthreads();
  }
}

// X = Y^T Y (X_ab = sum_{k >= max(a, b)} Y_ka Y_kb) for the lower Y; lower
// tiles computed, both triangles written.  256 threads x (4 x 4) per tile.
__global__ __launch_bounds__(256) void k_cov_gram(const double* __restrict__ Y, int n, double* __restrict__ X,
                                                  const LMState* st) {
  __shared__ double Ya[kCovK][kCovTile + 1];
  __shared__ double Yb[kCovK][kCovTile + 1];
  if (st->step_invalid) return;
  const int tt = blockIdx.x;
  int ta = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
  while (ta * (ta + 1) / 2 > tt) --ta;
  while ((ta + 1) * (ta + 2) / 2 <= tt) ++ta;
  const int tb = tt - ta * (ta + 1) / 2;
  const int a0 = ta * kCovTile, b0 = tb * kCovTile;
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  double acc[4][4] = {};
  // Y_ka = 0 for k < a, so rows below a0 (>= b0) contribute nothing
  for (int k0 = a0; k0 < n; k0 += kCovK) {
    for (int e = tid; e < kCovK * kCovTile; e += 256) {
      const int kk = e / kCovTile, c = e % kCovTile;
      const int k = k0 + kk;
      const int ca = a0 + c, cb = b0 + c;
      Ya[kk][c] = (k < n && ca < n && k >= ca) ? Y[(size_t)k * n + ca] : 0.0;
      Yb[kk][c] = (k < n && cb < n && k >= cb) ? Y[(size_t)k * n + cb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kCovK; ++kk) {
      double va[4], vb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) va[u] = Ya[kk][ty * 4 + u];
#pragma unroll
      for (int v = 0; v < 4; ++v) vb[v] = Yb[kk][tx * 4 + v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(va[u], vb[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = a0 + ty * 4 + u, b = b0 + tx * 4 + v;
      if (a < n && b < n) {
        X[(size_t)a * n + b] = acc[u][v];
        X[(size_t)b * n + a] = acc[u][v];
      }
    }
}

// Rank test on the pivots.  A Cholesky of J^T J cannot resolve eigenvalues
// below about eps ||J^T J||, so a pivot d_k = L_kk^2 below rtol times the
// largest pivot of its factor counts as a null direction (rtol is the
// reference's singular-value threshold 20 (m + n) eps, covariance.py:374-377,
// applied to pivots).  Coordinates of constant parameters (unit diagonal,
// decoupled) are not counted.  Every deficient pivot adds one to counts[0];
// a singular H also sets state->step_invalid so the later kernels skip.
__global__ __launch_bounds__(kPointThreads) void k_cov_check_points(const double* __restrict__ Linv,
                                                                    const uint8_t* __restrict__ point_mask, int P,
                                                                    double rtol, int32_t* counts, LMState* st) {
  const int p = blockIdx.x * kPointThreads + threadIdx.x;
  if (p >= P) return;
  const double* Li = Linv + 6 * (size_t)p;
  const double inv[3] = {Li[0], Li[2], Li[5]};  // 1 / l_kk
  double piv[3], top = 0.0;
  bool finite = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    piv[k] = 1.0 / (inv[k] * inv[k]);
    finite = finite && isfinite(inv[k]) && inv[k] != 0.0;
    const bool free_k = !point_mask || point_mask[3 * p + k];
    if (free_k && isfinite(piv[k])) top = fmax(top, piv[k]);
  }
  int bad = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool free_k = !point_mask || point_mask[3 * p + k];
    if (free_k && (!finite || !(piv[k] > rtol * top))) bad += 1;
  }
  if (bad) {
    atomicAdd(counts, bad);
    st->step_invalid = 1;
  }
}

constexpr int kPivotThreads = 1024;

__global__ __launch_bounds__(kPivotThreads) void k_cov_check_pivots(const double* __restrict__ L, int n,
                                                                    const uint8_t* __restrict__ camera_mask,
                                                                    double rtol, int32_t* counts, LMState* st) {
  __shared__ double red[kPivotThreads / kWave];
  double top = 0.0;
  for (int i = threadIdx.x; i < n; i += kPivotThreads) {
    const double d = L[(size_t)i * n + i];
    if ((!camera_mask || camera_mask[i]) && isfinite(d)) top = fmax(top, d * d);
  }
  top = block_max<kPivotThreads>(top, red);
  __shared__ double s_top;
  if (threadIdx.x == 0) s_top = top;
  __syncthreads();
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += kPivotThreads) {
    const double d = L[(size_t)i * n + i];
    if ((!camera_mask || camera_mask[i]) && !(d * d > rtol * s_top)) bad += 1;
  }
  if (bad) {
    atomicAdd(counts + 1, bad);
    st->step_invalid = 1;
  }
}

// Rank of the reduced camera system by Cholesky with symmetric pivoting, on
// S itself (both triangles, overwritten): at each step the largest remaining
// diagonal entry of the Schur complement is eliminated; the rank is the
// number of steps before that entry falls to rtol times the largest initial
// diagonal (the usual rank-revealing stop).  Coordinates of constant
// parameters are left out.  counts[1] = free coordinates - rank.  Only run
// when the unpivoted factorisation found a singular S, to count its null
// directions (an unpivoted factor's later pivots are meaningless after a
// rounding-level one).  One workgroup; remaining indices compacted in LDS.
constexpr int kRankThreads = 1024;

__global__ __launch_bounds__(kRankThreads) void k_cov_rank_pivoted(double* __restrict__ A, int n,
                                                                   const uint8_t* __restrict__ camera_mask,
                                                                   double rtol, int32_t* counts) {
  extern __shared__ int rem[];  // [n] remaining free coordinates
  __shared__ double red_v[kRankThreads / kWave];
  __shared__ int red_i[kRankThreads / kWave];
  __shared__ int s_count, s_piv;
  __shared__ double s_d, s_top;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  if (tid == 0) {
    int c = 0;
    double top = 0.0;
    for (int i = 0; i < n; ++i)
      if (!camera_mask || camera_mask[i]) {
        rem[c++] = i;
        const double d = A[(size_t)i * n + i];
        if (isfinite(d)) top = fmax(top, d);
      }
    s_count = c;
    s_top = top;
  }
  __syncthreads();
  const int free_count = s_count;
  int rank = 0;
  while (true) {
    const int m = s_count;
    // argmax of the remaining diagonal (ties: the smaller position)
    double best = -1.0;
    int bi = -1;
    for (int t = tid; t < m; t += kRankThreads) {
      const int i = rem[t];
      const double d = A[(size_t)i * n + i];
      if (d > best) {
        best = d;
        bi = t;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off, kWave);
      const int oi = __shfl_xor(bi, off, kWave);
      if (ov > best || (ov == best && oi >= 0 && (bi < 0 || oi < bi))) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[wid] = best;
      red_i[wid] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = -1.0;
      int bt = -1;
      for (int w = 0; w < kRankThreads / kWave; ++w)
        if (red_v[w] > bv || (red_v[w] == bv && red_i[w] >= 0 && (bt < 0 || red_i[w] < bt))) {
          bv = red_v[w];
          bt = red_i[w];
        }
      s_d = bv;
      s_piv = bt;
    }
    __syncthreads();
    const double d = s_d;
    const int pt = s_piv;
    if (m == 0 || pt < 0 || !(d > rtol * s_top) || !isfinite(d)) break;
    const int p = rem[pt];
    __syncthreads();
    if (tid == 0) {  // drop the pivot from the remaining list
      rem[pt] = rem[m - 1];
      s_count = m - 1;
    }
    __syncthreads();
    const int m1 = m - 1;
    // A_ij -= A_ip A_jp / d over the remaining (i, j), both triangles
    const double inv_d = 1.0 / d;
    for (size_t e = tid; e < (size_t)m1 * m1; e += kRankThreads) {
      const int i = rem[e / m1], j = rem[e % m1];
      A[(size_t)i * n + j] -= A[(size_t)i * n + p] * A[(size_t)j * n + p] * inv_d;
    }
    __syncthreads();
    ++rank;
  }
  if (tid == 0) counts[1] = free_count - rank;
}

__global__ void k_cov_clear(int32_t* counts) {
  if (threadIdx.x < 2) counts[threadIdx.x] = 0;
}

// (L_p^-1)_{s r} from the packed lower factor inverse (row-major lower triangle)
__device__ __forceinline__ double linv_at(const double* Li, int s, int r) {
  return r <= s ? Li[s * (s + 1) / 2 + r] : 0.0;
}

// Lanes e < 27 write F_o[e / 9][e % 9] = (L_p^-T P_o^T Jc_o)_{rm} into F.
__device__ __forceinline__ void stage_f(const double* __restrict__ JP, const double* __restrict__ Li, int o,
                                        double* F) {
  const int e = threadIdx.x;
  if (e < 27) {
    const int r = e / 9, m = e % 9;
    const double* rec = JP + (size_t)kJPStride * o;
    double v = 0.0;
    for (int s = r; s < 3; ++s) {
      const double pj = rec[18 + s] * rec[m] + rec[21 + s] * rec[9 + m];  // (P^T Jc)_{s m}
      v = fma(linv_at(Li, s, r), pj, v);
    }
    F[e] = v;
  }
}

struct CovArgs {
  const double* X;              // [n][n] S^-1
  int n;
  const double* JP;             // [N][kJPStride] camera Jacobian and P per observation
  const double* Linv;           // [P][6] point factor inverses
  const int32_t* obs_camera;    // [N]
  const int32_t* point_offsets; // [P+1]
  const int32_t* requests;      // [R][4] kind a, index a, kind b, index b (0 camera, 1 point)
  double* out;                  // [R][81] each block row-major (rows x cols of its two blocks)
  const LMState* st;
};

// One wavefront per requested block.
__global__ __launch_bounds__(kWave) void k_cov_blocks(CovArgs a) {
  __shared__ double F0[27], F1[27], T[81], M[27];
  if (a.st->step_invalid) return;
  const int lane = threadIdx.x;
  const int32_t* q = a.requests + 4 * (size_t)blockIdx.x;
  int ka = q[0], ia = q[1], kb = q[2], ib = q[3];
  double* out = a.out + 81 * (size_t)blockIdx.x;
  const int n = a.n;
  if (ka == 0 && kb == 0) {  // (S^-1)_{ia ib}
    for (int e = lane; e < 81; e += kWave) out[e] = a.X[(size_t)(9 * ia + e / 9) * n + 9 * ib + e % 9];
    return;
  }
  bool transpose = false;
  if (ka == 0) {  // camera-point: the transpose of point-camera
    transpose = true;
    int t = ka; ka = kb; kb = t;
    t = ia; ia = ib; ib = t;
  }
  if (kb == 0) {  // point ia, camera ib: -sum_o F_o (S^-1)_{c(o) ib}, 3 x 9
    const double* Li = a.Linv + 6 * (size_t)ia;
    const int r = lane / 9, c = lane % 9;
    double acc = 0.0;
    for (int o = a.point_offsets[ia]; o < a.point_offsets[ia + 1]; ++o) {
      stage_f(a.JP, Li, o, F0);
      __syncthreads();
      const int co = a.obs_camera[o];
      if (lane < 27)
        for (int m = 0; m < 9; ++m) acc = fma(F0[r * 9 + m], a.X[(size_t)(9 * co + m) * n + 9 * ib + c], acc);
      __syncthreads();
    }
    if (lane < 27) {
      if (transpose) out[c * 3 + r] = -acc;
      else out[r * 9 + c] = -acc;
    }
    return;
  }
  // point ia, point ib: [ia = ib] V^-1 + sum_{o, o'} F_o (S^-1)_{c(o) c(o')} F_o'^T, 3 x 3
  const double* La = a.Linv + 6 * (size_t)ia;
  const double* Lb = a.Linv + 6 * (size_t)ib;
  const int r = lane / 3, c = lane % 3;
  double acc = 0.0;
  for (int o = a.point_offsets[ia]; o < a.point_offsets[ia + 1]; ++o) {
    stage_f(a.JP, La, o, F0);
    const int co = a.obs_camera[o];
    for (int o2 = a.point_offsets[ib]; o2 < a.point_offsets[ib + 1]; ++o2) {
      stage_f(a.JP, Lb, o2, F1);
      const int co2 = a.obs_camera[o2];
      for (int e = lane; e < 81; e += kWave) T[e] = a.X[(size_t)(9 * co + e / 9) * n + 9 * co2 + e % 9];
      __syncthreads();
      if (lane < 27) {  // M = F_o T, 3 x 9
        const int mr = lane / 9, mc = lane % 9;
        double v = 0.0;
        for (int m = 0; m < 9; ++m) v = fma(F0[mr * 9 + m], T[m * 9 + mc], v);
        M[lane] = v;
      }
      __syncthreads();
      if (lane < 9)
        for (int m = 0; m < 9; ++m) acc = fma(M[r * 9 + m], F1[c * 9 + m], acc);
      __syncthreads();
    }
  }
  if (lane < 9) {
    if (ia == ib) {
      double v = 0.0;
      for (int s = (r > c ? r : c); s < 3; ++s) v = fma(linv_at(La, s, r), linv_at(La, s, c), v);
      acc += v;
    }
    out[r * 3 + c] = acc;
  }
}

}  // namespace

size_t covariance_work_doubles(int n) { return 2 * (size_t)n * (size_t)n; }

int launch_covariance(const CovLaunch& c, hipStream_t s) {
  hipLaunchKernelGGL(k_cov_clear, dim3(1), dim3(64), 0, s, c.counts);
  CT_LAUNCH_CHECK();
  if (c.num_points > 0) {
    hipLaunchKernelGGL(k_cov_check_points, dim3((c.num_points + kPointThreads - 1) / kPointThreads),
                       dim3(kPointThreads), 0, s, c.Linv, c.point_mask, c.num_points, c.pivot_rtol, c.counts, c.st);
    CT_LAUNCH_CHECK();
  }
  const int n = c.n;
  double* Y = c.work;
  double* X = c.work + (size_t)n * n;
  if (n > 0) {
    hipLaunchKernelGGL(k_cov_check_pivots, dim3(1), dim3(kPivotThreads), 0, s, c.L, n, c.camera_mask, c.pivot_rtol,
                       c.counts, c.st);
    CT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cov_lower_inverse, dim3(n), dim3(kWave), sizeof(double) * (size_t)n, s, c.L, n, Y, c.st);
    CT_LAUNCH_CHECK();
    const int T = (n + kCovTile - 1) / kCovTile;
    hipLaunchKernelGGL(k_cov_gram, dim3(T * (T + 1) / 2), dim3(256), 0, s, Y, n, X, c.st);
    CT_LAUNCH_CHECK();
  }
  if (c.num_requests > 0) {
    CovArgs a{X, n, c.JP, c.Linv, c.obs_camera, c.point_offsets, c.requests, c.out, c.st};
    hipLaunchKernelGGL(k_cov_blocks, dim3((unsigned int)c.num_requests), dim3(kWave), 0, s, a);
    CT_LAUNCH_CHECK();
  }
  return CT_OK;
}

int launch_covariance_rank(double* S, int n, const uint8_t* camera_mask, double pivot_rtol, int32_t* counts,
                           hipStream_t s) {
  if (n <= 0) return CT_OK;
  hipLaunchKernelGGL(k_cov_rank_pivoted, dim3(1), dim3(kRankThreads), sizeof(int) * (size_t)n, s, S, n, camera_mask,
                     pivot_rtol, counts);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct
