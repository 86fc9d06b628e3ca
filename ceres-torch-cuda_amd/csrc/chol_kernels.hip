// FP64 solves of the reduced camera system S x = rhs (S symmetric positive
// definite, n x n row-major, lower triangle referenced).
//
//  * n <= 256: one workgroup factors S blockwise (16-column panels: the
//    diagonal block is factored in LDS by all threads, the rows below are
//    solved one row per thread and the trailing rank-16 update runs on 4x4
//    register tiles) and runs both triangular solves.
//  * n > 256: right-looking blocked Cholesky, 32-column panels; per panel one
//    wavefront factors the diagonal block, one kernel solves the rows below it
//    and one applies the trailing update on 64x64 tiles; the triangular solves
//    run in one persistent many-workgroup launch after the 64x64 diagonal
//    blocks have been inverted.
//  * explicit-S PCG with 9x9 block-Jacobi (SCHUR_JACOBI) preconditioning as
//    one persistent kernel (linear.py:1216-1256, preconditioner :1156-1182).
#include "ba_kernels.h"

namespace ct {

namespace {

__device__ __forceinline__ bool skip_solve(const LMState* st) { return st != nullptr && (st->done || st->skip_step); }

// Factor the kb x kb lower block held in LDS (row stride NBS) in place with
// every thread of the workgroup: per column, all threads read the pivot, the
// column below it is scaled by 1/sqrt(pivot) and each trailing lower entry
// gets its rank-1 update from one thread.  dinv[j] = 1/L[j][j].  A pivot that
// is not positive and finite sets *bad.
template <int NB, int NBS, int THREADS>
__device__ void factor_diag_block(double (*Ld)[NBS], double* dinv, int kb, int* bad) {
  const int tid = threadIdx.x;
  for (int j = 0; j < kb; ++j) {
    const double piv = Ld[j][j];
    const double d = sqrt(piv);
    const double inv = 1.0 / d;
    __syncthreads();  // the pivot is read by all threads before it is replaced
    if (tid == 0) {
      if (!(piv > 0.0 && isfinite(piv))) *bad = 1;
      Ld[j][j] = d;
      dinv[j] = inv;
    }
    for (int i = j + 1 + tid; i < kb; i += THREADS) Ld[i][j] *= inv;
    __syncthreads();
    for (int idx = tid; idx < NB * NB; idx += THREADS) {
      const int i = idx / NB, k = idx % NB;
      if (k > j && k <= i && i < kb) Ld[i][k] -= Ld[i][j] * Ld[k][j];
    }
    __syncthreads();
  }
}

// One row of the panel solve v = a L11^-T, i.e.
//   v[c] = (a[c] - sum_{l<c} v[l] L[c][l]) / L[c][c],
// in chunks of 8 columns held in registers.  prow (the thread's own LDS row)
// holds a on entry and v on exit.  Ld rows >= kb must be zero.
template <int NB, int NBS>
__device__ __forceinline__ void trsm_row(const double (*Ld)[NBS], const double* dinv, int kb, double* prow) {
  static_assert(NB % 8 == 0, "panel width must be a multiple of 8");
  for (int c0 = 0; c0 < kb; c0 += 8) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = c0 + k < kb ? prow[c0 + k] : 0.0;
    for (int l = 0; l < c0; ++l) {
      const double pl = prow[l];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] -= pl * Ld[c0 + k][l];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double s = v[k];
#pragma unroll
      for (int l = 0; l < k; ++l) s -= v[l] * Ld[c0 + k][c0 + l];
      v[k] = c0 + k < kb ? s * dinv[c0 + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c0 + k < kb) prow[c0 + k] = v[k];
  }
}

// Forward substitution of one kb-row block, wave 0 only (lane l owns row l):
// solves L[k0:k0+kb, k0:k0+kb] yb = y[k0:k0+kb] with the pivots' reciprocals.
__device__ __forceinline__ double tri_forward_block(const double* __restrict__ A, const double* __restrict__ dinv,
                                                    int n, int k0, int kb, double yl) {
  const int l = threadIdx.x;
  for (int j = 0; j < kb; ++j) {
    const double yj = __shfl(yl, j) * dinv[k0 + j];
    if (l == j) yl = yj;
    else if (l > j && l < kb) yl -= A[(size_t)(k0 + l) * n + k0 + j] * yj;
  }
  return yl;
}

// Backward substitution of one block with L^T, wave 0 only.
__device__ __forceinline__ double tri_backward_block(const double* __restrict__ A, const double* __restrict__ dinv,
                                                     int n, int k0, int kb, double yl) {
  const int l = threadIdx.x;
  for (int j = kb - 1; j >= 0; --j) {
    const double yj = __shfl(yl, j) * dinv[k0 + j];
    if (l == j) yl = yj;
    else if (l < j) yl -= A[(size_t)(k0 + j) * n + k0 + l] * yj;
  }
  return yl;
}

// ---------------------------------------------------------------------------
// small n: one workgroup
// ---------------------------------------------------------------------------
constexpr int kSmallNB = 16;
constexpr int kSmallMax = 256;
constexpr int kSmallThreads = 256;

__global__ __launch_bounds__(kSmallThreads) void k_chol_solve_small(double* __restrict__ A,
                                                                    const double* __restrict__ rhs,
                                                                    double* __restrict__ x, double* __restrict__ tmp,
                                                                    int n, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double Ld[kSmallNB][kSmallNB + 1];
  __shared__ double Pn[kSmallMax][kSmallNB + 1];
  __shared__ double dinv[kSmallMax];
  __shared__ double y[kSmallMax];
  __shared__ double yb[kSmallNB];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += kSmallNB) {
    const int kb = min(kSmallNB, n - k0);
    for (int idx = tid; idx < kSmallNB * kSmallNB; idx += kSmallThreads) {
      const int r = idx / kSmallNB, c = idx % kSmallNB;
      Ld[r][c] = (r < kb && c <= r) ? A[(size_t)(k0 + r) * n + k0 + c] : 0.0;
    }
    __syncthreads();
    factor_diag_block<kSmallNB, kSmallNB + 1, kSmallThreads>(Ld, dinv + k0, kb, &bad);
    for (int idx = tid; idx < kSmallNB * kSmallNB; idx += kSmallThreads) {
      const int r = idx / kSmallNB, c = idx % kSmallNB;
      if (r < kb && c <= r) A[(size_t)(k0 + r) * n + k0 + c] = Ld[r][c];
    }
    // rows below the block: L21 = A21 L11^-T
    const int rows = n - k0 - kb;
    for (int i = tid; i < rows; i += kSmallThreads) {
      const int row = k0 + kb + i;
      for (int c = 0; c < kb; ++c) Pn[i][c] = A[(size_t)row * n + k0 + c];
      trsm_row<kSmallNB, kSmallNB + 1>(Ld, dinv + k0, kb, Pn[i]);
      for (int c = 0; c < kb; ++c) A[(size_t)row * n + k0 + c] = Pn[i][c];
    }
    __syncthreads();
    // trailing update A22 -= L21 L21^T on 4x4 register tiles of the lower triangle
    const int T = (rows + 3) / 4;
    const int ntiles = T * (T + 1) / 2;
    for (int tt = tid; tt < ntiles; tt += kSmallThreads) {
      int ti = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
      while (ti * (ti + 1) / 2 > tt) --ti;
      while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
      const int tj = tt - ti * (ti + 1) / 2;
      double acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
      for (int c = 0; c < kb; ++c) {
        double pa[4], pb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int i = 4 * ti + a, j = 4 * tj + a;
          pa[a] = i < rows ? Pn[i][c] : 0.0;
          pb[a] = j < rows ? Pn[j][c] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] += pa[a] * pb[b];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int i = 4 * ti + a, j = 4 * tj + b;
          if (i < rows && j <= i) A[(size_t)(k0 + kb + i) * n + k0 + kb + j] -= acc[a][b];
        }
      }
    }
    __syncthreads();
  }
  // keep the pivots' reciprocals for later triangular solves with this factor
  for (int i = tid; i < n; i += kSmallThreads) {
    tmp[n + i] = dinv[i];
    y[i] = rhs ? rhs[i] : 0.0;
  }
  __syncthreads();
  if (rhs == nullptr) {  // factorisation only
    if (tid == 0 && st && bad) st->step_invalid = 1;
    return;
  }
  // forward substitution L y = rhs
  for (int k0 = 0; k0 < n; k0 += kSmallNB) {
    const int kb = min(kSmallNB, n - k0);
    if (tid < kWave) {
      const double yl = tri_forward_block(A, dinv, n, k0, kb, tid < kb ? y[k0 + tid] : 0.0);
      if (tid < kb) {
        y[k0 + tid] = yl;
        yb[tid] = yl;
      }
    }
    __syncthreads();
    for (int i = k0 + kb + tid; i < n; i += kSmallThreads) {
      double s = 0.0;
      for (int c = 0; c < kb; ++c) s += A[(size_t)i * n + k0 + c] * yb[c];
      y[i] -= s;
    }
    __syncthreads();
  }
  // backward substitution L^T x = y
  const int last = ((n - 1) / kSmallNB) * kSmallNB;
  for (int k0 = last; k0 >= 0; k0 -= kSmallNB) {
    const int kb = min(kSmallNB, n - k0);
    if (tid < kWave) {
      const double yl = tri_backward_block(A, dinv, n, k0, kb, tid < kb ? y[k0 + tid] : 0.0);
      if (tid < kb) {
        y[k0 + tid] = yl;
        yb[tid] = yl;
      }
    }
    __syncthreads();
    for (int i = tid; i < k0; i += kSmallThreads) {
      double s = 0.0;
      for (int c = 0; c < kb; ++c) s += A[(size_t)(k0 + c) * n + i] * yb[c];
      y[i] -= s;
    }
    __syncthreads();
  }
  int my_bad = 0;
  for (int i = tid; i < n; i += kSmallThreads) {
    x[i] = y[i];
    if (!isfinite(y[i])) my_bad = 1;
  }
  if (my_bad) bad = 1;
  __syncthreads();
  if (tid == 0 && st) {
    if (bad) st->step_invalid = 1;
    st->linear_iterations = 1;
  }
}

// ---------------------------------------------------------------------------
// large n: blocked right-looking factorisation
// ---------------------------------------------------------------------------
constexpr int kNB = 32;
constexpr int kPanelThreads = 256;
constexpr int kTile = 64;

__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Factor the diagonal block of panel k0 in place with one wavefront: lane l
// holds row l in registers; column j is broadcast through LDS.  Stores the
// pivots' reciprocals.
__global__ __launch_bounds__(kWave) void k_chol_diag(double* __restrict__ A, double* __restrict__ dinv_out, int n,
                                                     int k0, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double col[kNB];
  const int l = threadIdx.x;
  const int kb = min(kNB, n - k0);
  const bool own = l < kb;
  double a[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) a[c] = (own && c <= l) ? A[(size_t)(k0 + l) * n + k0 + c] : 0.0;
  int bad = 0;
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    if (j < kb) {
      const double piv = __shfl(a[j], j);
      if (!(piv > 0.0) || !isfinite(piv)) bad = 1;
      const double d = sqrt(piv);
      const double inv = 1.0 / d;
      const double lj = (l == j) ? d : a[j] * inv;
      if (l >= j) a[j] = lj;
      if (l < kNB) col[l] = (l >= j && own) ? lj : 0.0;
      lds_wave_sync();
#pragma unroll
      for (int m = j + 1; m < kNB; ++m)
        if (m <= l) a[m] -= lj * col[m];
      if (l == j && own) dinv_out[k0 + j] = inv;
      lds_wave_sync();
    }
  }
  if (own) {
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (c <= l) A[(size_t)(k0 + l) * n + k0 + c] = a[c];
  }
  if (l == 0 && bad && st) st->step_invalid = 1;
}

// Rows below the factored diagonal block of panel k0: L21 = A21 L11^-T, one
// row per thread.
__global__ __launch_bounds__(kPanelThreads) void k_chol_panel(double* __restrict__ A,
                                                              const double* __restrict__ dinv_in, int n, int k0,
                                                              LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double Ld[kNB][kNB + 1];
  __shared__ double P[kPanelThreads][kNB + 1];
  __shared__ double dinv[kNB];
  const int tid = threadIdx.x;
  const int kb = min(kNB, n - k0);
  for (int idx = tid; idx < kNB * kNB; idx += kPanelThreads) {
    const int r = idx / kNB, c = idx % kNB;
    Ld[r][c] = (r < kb && c <= r) ? A[(size_t)(k0 + r) * n + k0 + c] : 0.0;
  }
  if (tid < kNB) dinv[tid] = tid < kb ? dinv_in[k0 + tid] : 0.0;
  __syncthreads();
  const int row = k0 + kb + blockIdx.x * kPanelThreads + tid;
  if (row < n) {
    double* prow = P[tid];
    for (int c = 0; c < kb; ++c) prow[c] = A[(size_t)row * n + k0 + c];
    trsm_row<kNB, kNB + 1>(Ld, dinv, kb, prow);
    for (int c = 0; c < kb; ++c) A[(size_t)row * n + k0 + c] = prow[c];
  }
}

// A[i][j] -= sum_c L[i][k0+c] L[j][k0+c] over 64x64 lower tiles of the
// trailing matrix (rows/cols >= r0 = k0 + kb).  256 threads x (4x4).
__global__ __launch_bounds__(256) void k_chol_syrk(double* __restrict__ A, int n, int k0, int kb, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double Li[kTile][kNB + 1];
  __shared__ double Lj[kTile][kNB + 1];
  const int r0 = k0 + kb;
  const int tt = blockIdx.x;
  int ti = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > tt) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
  const int tj = tt - ti * (ti + 1) / 2;
  const int i0 = r0 + ti * kTile, j0 = r0 + tj * kTile;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kTile * kNB; idx += 256) {
    const int r = idx / kNB, c = idx % kNB;
    const int gi = i0 + r, gj = j0 + r;
    Li[r][c] = (gi < n && c < kb) ? A[(size_t)gi * n + k0 + c] : 0.0;
    Lj[r][c] = (gj < n && c < kb) ? A[(size_t)gj * n + k0 + c] : 0.0;
  }
  __syncthreads();
  const int tr = tid / 16, tc = tid % 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int c = 0; c < kb; ++c) {
    double pa[4], pb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      pa[a] = Li[tr + 16 * a][c];
      pb[a] = Lj[tc + 16 * a][c];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += pa[a] * pb[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gi = i0 + tr + 16 * a, gj = j0 + tc + 16 * b;
      if (gi < n && gj <= gi) A[(size_t)gi * n + gj] -= acc[a][b];
    }
  }
}

// Blocked triangular solves with one workgroup against the factor in A and
// the pivots' reciprocals tmp[n:2n]; y lives in tmp[0:n].
constexpr int kTriThreads = 1024;
__global__ __launch_bounds__(kTriThreads) void k_tri_solve(const double* __restrict__ A,
                                                            const double* __restrict__ rhs, double* __restrict__ x,
                                                            double* __restrict__ tmp, int n, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double yb[kSmallNB];
  __shared__ int bad;
  double* __restrict__ y = tmp;
  const double* __restrict__ dinv = tmp + n;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (int i = tid; i < n; i += kTriThreads) y[i] = rhs[i];
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += kSmallNB) {
    const int kb = min(kSmallNB, n - k0);
    if (tid < kWave) {
      const double yl = tri_forward_block(A, dinv, n, k0, kb, tid < kb ? y[k0 + tid] : 0.0);
      if (tid < kb) {
        y[k0 + tid] = yl;
        yb[tid] = yl;
      }
    }
    __syncthreads();
    for (int i = k0 + kb + tid; i < n; i += kTriThreads) {
      double s = 0.0;
      for (int c = 0; c < kb; ++c) s += A[(size_t)i * n + k0 + c] * yb[c];
      y[i] -= s;
    }
    __syncthreads();
  }
  const int last = ((n - 1) / kSmallNB) * kSmallNB;
  for (int k0 = last; k0 >= 0; k0 -= kSmallNB) {
    const int kb = min(kSmallNB, n - k0);
    if (tid < kWave) {
      const double yl = tri_backward_block(A, dinv, n, k0, kb, tid < kb ? y[k0 + tid] : 0.0);
      if (tid < kb) {
        y[k0 + tid] = yl;
        yb[tid] = yl;
      }
    }
    __syncthreads();
    for (int i = tid; i < k0; i += kTriThreads) {
      double s = 0.0;
      for (int c = 0; c < kb; ++c) s += A[(size_t)(k0 + c) * n + i] * yb[c];
      y[i] -= s;
    }
    __syncthreads();
  }
  int my_bad = 0;
  for (int i = tid; i < n; i += kTriThreads) {
    x[i] = y[i];
    if (!isfinite(y[i])) my_bad = 1;
  }
  if (my_bad) bad = 1;
  __syncthreads();
  if (tid == 0 && st) {
    if (bad) st->step_invalid = 1;
    st->linear_iterations = 1;
  }
}

// ---------------------------------------------------------------------------
// large n: triangular solves with many workgroups
//
// The factor is cut into kTB-row blocks.  k_diag_inverse inverts every
// diagonal block (one workgroup each, all at once).  k_tri_solve_mt then
// solves L y = b and L^T x = y in one persistent launch: workgroup g owns
// forward blocks g, g+G, ... (increasing) and backward blocks nt-1-g,
// nt-1-g-G, ... (decreasing).  Block r of the forward solve subtracts
// L_rj y_j as each y_j is published and finishes with y_r = L_rr^-1 (...);
// the backward solve mirrors it with L^T.  Publication is a per-block flag
// stamped with this launch's epoch (read once at start; the last workgroup
// to finish advances it), so flags never need clearing.  Each workgroup
// waits only on blocks that precede its own in the dependency order and all
// G <= #CU workgroups are resident, so every wait ends; waits are bounded
// anyway and a timeout flags the step invalid.
// ---------------------------------------------------------------------------
constexpr int kTB = 64;
constexpr int kTriMtThreads = 256;
constexpr int kTriSlices = kTriMtThreads / kTB;  // column slices per row in a block product

__host__ __device__ inline size_t tri_blocks(int n) { return ((size_t)n + kTB - 1) / kTB; }

// inverse of the lower-triangular diagonal block b of the factor in A, row-major
// kTB x kTB (zero outside the block's rows), one lane per column
__global__ __launch_bounds__(kWave) void k_diag_inverse(const double* __restrict__ A, double* __restrict__ linv,
                                                        int n, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double Ls[kTB][kTB + 1];
  __shared__ double dinv[kTB];
  const int b = blockIdx.x;
  const int r0 = b * kTB;
  const int rb = min(kTB, n - r0);
  const int c = threadIdx.x;
  for (int i = 0; i < kTB; ++i) Ls[i][c] = (i < rb && c <= i) ? A[(size_t)(r0 + i) * n + r0 + c] : 0.0;
  lds_wave_sync();
  dinv[c] = c < rb ? 1.0 / Ls[c][c] : 0.0;
  lds_wave_sync();
  double x[kTB];
#pragma unroll
  for (int i = 0; i < kTB; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) s += Ls[i][k] * x[k];
    x[i] = (i < c) ? 0.0 : (i == c ? dinv[i] : -s * dinv[i]);
  }
  double* out = linv + (size_t)b * kTB * kTB;
#pragma unroll
  for (int i = 0; i < kTB; ++i) out[i * kTB + c] = (i < rb && c < rb) ? x[i] : 0.0;
}

struct TriWork {
  double* y;             // [n] forward solution
  const double* linv;    // [nt][kTB][kTB] inverses of the diagonal blocks
  int* flags;            // [2 nt] epoch stamps: y block ready, x block ready
  unsigned int* epoch;   // [2] completed launches, finished workgroups of this launch
};

__device__ __forceinline__ bool wait_stamp(const int* flag, int epoch, int* timed_out) {
  if (threadIdx.x == 0 && !*timed_out) {
    long spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
      if (++spins > (1L << 24)) {
        *timed_out = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
  return *timed_out == 0;
}

__device__ __forceinline__ void publish_stamp(int* flag, int epoch) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ __launch_bounds__(kTriMtThreads) void k_tri_solve_mt(const double* __restrict__ A,
                                                                const double* __restrict__ rhs,
                                                                double* __restrict__ x, int n, TriWork tw,
                                                                LMState* st, int G) {
  if (skip_solve(st)) return;
  __shared__ double acc[kTB];
  __shared__ double part[kTriSlices][kTB];
  __shared__ int timed_out;
  const int tid = threadIdx.x;
  const int row = tid % kTB, sl = tid / kTB;
  const int g = blockIdx.x;
  const int nt = (int)tri_blocks(n);
  const int epoch = (int)__hip_atomic_load(&tw.epoch[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
  if (tid == 0) timed_out = 0;
  __syncthreads();
  // forward: y_r = L_rr^-1 (b_r - sum_{j<r} L_rj y_j)
  for (int r = g; r < nt; r += G) {
    const int r0 = r * kTB, rb = min(kTB, n - r0);
    if (tid < kTB) acc[tid] = tid < rb ? rhs[r0 + tid] : 0.0;
    __syncthreads();
    for (int j = 0; j < r; ++j) {
      if (!wait_stamp(&tw.flags[j], epoch, &timed_out)) break;
      double p = 0.0;
      if (row < rb) {
        const double* Lrow = A + (size_t)(r0 + row) * n + j * kTB + sl * (kTB / kTriSlices);
        const double* yj = tw.y + j * kTB + sl * (kTB / kTriSlices);
#pragma unroll
        for (int c = 0; c < kTB / kTriSlices; ++c) p += Lrow[c] * yj[c];
      }
      part[sl][row] = p;
      __syncthreads();
      if (tid < kTB) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < kTriSlices; ++q) t += part[q][tid];
        acc[tid] -= t;
      }
      __syncthreads();
    }
    {
      const double* Li = tw.linv + (size_t)r * kTB * kTB + row * kTB + sl * (kTB / kTriSlices);
      double p = 0.0;
#pragma unroll
      for (int c = 0; c < kTB / kTriSlices; ++c) p += Li[c] * acc[sl * (kTB / kTriSlices) + c];
      part[sl][row] = p;
      __syncthreads();
      if (tid < rb) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < kTriSlices; ++q) t += part[q][tid];
        tw.y[r0 + tid] = t;
      }
    }
    publish_stamp(&tw.flags[r], epoch);
  }
  // backward: x_r = L_rr^-T (y_r - sum_{j>r} L_jr^T x_j)
  for (int q0 = g; q0 < nt; q0 += G) {
    const int r = nt - 1 - q0;
    const int r0 = r * kTB, rb = min(kTB, n - r0);
    if (!wait_stamp(&tw.flags[r], epoch, &timed_out)) break;
    if (tid < kTB) acc[tid] = tid < rb ? tw.y[r0 + tid] : 0.0;
    __syncthreads();
    for (int j = nt - 1; j > r; --j) {
      if (!wait_stamp(&tw.flags[nt + j], epoch, &timed_out)) break;
      const int j0 = j * kTB, jb = min(kTB, n - j0);
      double p = 0.0;
      if (row < rb) {
        for (int k = sl * (kTB / kTriSlices); k < (sl + 1) * (kTB / kTriSlices); ++k)
          if (k < jb) p += A[(size_t)(j0 + k) * n + r0 + row] * x[j0 + k];
      }
      part[sl][row] = p;
      __syncthreads();
      if (tid < kTB) {
        double t = 0.0;
#pragma unroll
        for (int qq = 0; qq < kTriSlices; ++qq) t += part[qq][tid];
        acc[tid] -= t;
      }
      __syncthreads();
    }
    {
      const double* Lb = tw.linv + (size_t)r * kTB * kTB;
      double p = 0.0;
      for (int c = sl * (kTB / kTriSlices); c < (sl + 1) * (kTB / kTriSlices); ++c) p += Lb[c * kTB + row] * acc[c];
      part[sl][row] = p;
      __syncthreads();
      if (tid < rb) {
        double t = 0.0;
#pragma unroll
        for (int qq = 0; qq < kTriSlices; ++qq) t += part[qq][tid];
        x[r0 + tid] = t;
      }
    }
    publish_stamp(&tw.flags[nt + r], epoch);
  }
  // non-finite solution or a timed-out wait invalidates the step
  int my_bad = 0;
  for (int q0 = g; q0 < nt; q0 += G) {
    const int r0 = (nt - 1 - q0) * kTB;
    if (tid < kTB && r0 + tid < n && !isfinite(x[r0 + tid])) my_bad = 1;
  }
  if (st && (my_bad || timed_out)) st->step_invalid = 1;
  if (st && g == 0 && tid == 0) st->linear_iterations = 1;
  // the last workgroup to finish advances the epoch for the next launch
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const unsigned int done = atomicAdd(&tw.epoch[1], 1u);
    if (done == (unsigned int)G - 1) {
      tw.epoch[1] = 0u;
      __threadfence();
      __hip_atomic_store(&tw.epoch[0], (unsigned int)epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---------------------------------------------------------------------------
// explicit-S PCG with the SCHUR_JACOBI preconditioner, one persistent kernel
// (linear.py:1216-1256; preconditioner linear.py:1156-1182, applied as in
// linear.py:69-77).  Workgroup g owns the 9-row camera blocks g, g+G, ...;
// its 9 wavefronts compute one row each of S p.  Every workgroup keeps the
// whole search direction p in LDS and updates it identically from the
// shared z, so an iteration needs two grid barriers (after p.Ap and after
// r.z).  Cross-workgroup sums read the per-workgroup partials in one fixed
// order in every workgroup, so all of them take identical decisions.  The
// grid barrier is a monotonic arrival counter (cleared by the LM begin
// kernel) with a bounded spin; on timeout the step is flagged invalid.
// ---------------------------------------------------------------------------
constexpr int kBlk = 9;
constexpr int kPcgWaves = kBlk;  // one wavefront per row of a camera block
constexpr int kPcgThreads = kPcgWaves * kWave;
constexpr int kPcgMaxOwn = 8;    // camera blocks per workgroup
constexpr int kPcgMaxGroups = 1024;
constexpr int kPcgMaxN = 12288;  // p is held in LDS
constexpr double kEps = 2.220446049250313e-16;

struct PcgWork {
  double* z;          // [n]  preconditioned residual, read by every workgroup
  double* part;       // [3][kPcgMaxGroups] partials: r.r, r.z, p.Ap
  unsigned int* bar;  // arrival counter
};

__device__ __forceinline__ bool grid_barrier(unsigned int* bar, unsigned int target, int* timed_out) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(bar, 1u);
    long spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > (1L << 26)) {
        *timed_out = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  __syncthreads();
  return *timed_out == 0;
}

// Sum of part[0:G] in one fixed order (identical in every workgroup); valid
// in every thread after the call.
__device__ __forceinline__ double group_total(const double* part, unsigned int G, double* slot) {
  if (threadIdx.x < kWave) {
    double v = 0.0;
    for (unsigned int i = threadIdx.x; i < G; i += kWave) v += part[i];
    v = wave_sum(v);
    if (threadIdx.x == 0) *slot = v;
  }
  __syncthreads();
  const double t = *slot;
  __syncthreads();
  return t;
}

// Cholesky of a 9x9 block in place (lower triangle), reciprocal pivots in dinv.
__device__ __forceinline__ bool chol9(double* A, double* dinv) {
  for (int j = 0; j < kBlk; ++j) {
    double d = A[j * kBlk + j];
    for (int k = 0; k < j; ++k) d -= A[j * kBlk + k] * A[j * kBlk + k];
    if (!(d > 0.0) || !isfinite(d)) return false;
    d = sqrt(d);
    const double inv = 1.0 / d;
    A[j * kBlk + j] = d;
    dinv[j] = inv;
    for (int i = j + 1; i < kBlk; ++i) {
      double t = A[i * kBlk + j];
      for (int k = 0; k < j; ++k) t -= A[i * kBlk + k] * A[j * kBlk + k];
      A[i * kBlk + j] = t * inv;
    }
  }
  return true;
}

// LU with partial pivoting of a 9x9 block in place (torch.linalg.solve's
// factorisation); reciprocal pivots in dinv.  False if a pivot is zero.
__device__ __forceinline__ bool lu9(double* A, int* piv, double* dinv) {
  for (int j = 0; j < kBlk; ++j) {
    int p = j;
    double best = fabs(A[j * kBlk + j]);
    for (int i = j + 1; i < kBlk; ++i) {
      const double v = fabs(A[i * kBlk + j]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[j] = p;
    if (!(best > 0.0) || !isfinite(best)) return false;
    if (p != j)
      for (int k = 0; k < kBlk; ++k) {
        const double t = A[j * kBlk + k];
        A[j * kBlk + k] = A[p * kBlk + k];
        A[p * kBlk + k] = t;
      }
    const double inv = 1.0 / A[j * kBlk + j];
    dinv[j] = inv;
    for (int i = j + 1; i < kBlk; ++i) {
      const double l = A[i * kBlk + j] * inv;
      A[i * kBlk + j] = l;
      for (int k = j + 1; k < kBlk; ++k) A[i * kBlk + k] -= l * A[j * kBlk + k];
    }
  }
  return true;
}

// z = B^-1 r for one block: Cholesky solve (kind 1, in registers) or LU
// solve (kind 0, rare: worked in place in z, which must not alias r).
__device__ __forceinline__ void block_solve(const double* A, const double* dinv, const int* piv, int kind, const double* r,
                            double* z) {
  if (kind == 1) {
    double t[kBlk];
#pragma unroll
    for (int i = 0; i < kBlk; ++i) {
      double s = r[i];
#pragma unroll
      for (int k = 0; k < i; ++k) s -= A[i * kBlk + k] * t[k];
      t[i] = s * dinv[i];
    }
#pragma unroll
    for (int i = kBlk - 1; i >= 0; --i) {
      double s = t[i];
#pragma unroll
      for (int k = i + 1; k < kBlk; ++k) s -= A[k * kBlk + i] * t[k];
      t[i] = s * dinv[i];
    }
#pragma unroll
    for (int i = 0; i < kBlk; ++i) z[i] = t[i];
    return;
  }
  for (int i = 0; i < kBlk; ++i) z[i] = r[i];
  for (int j = 0; j < kBlk; ++j) {
    const int p = piv[j];
    const double t = z[j];
    z[j] = z[p];
    z[p] = t;
  }
  for (int i = 0; i < kBlk; ++i) {
    double s = z[i];
    for (int k = 0; k < i; ++k) s -= A[i * kBlk + k] * z[k];
    z[i] = s;
  }
  for (int i = kBlk - 1; i >= 0; --i) {
    double s = z[i];
    for (int k = i + 1; k < kBlk; ++k) s -= A[i * kBlk + k] * z[k];
    z[i] = s * dinv[i];
  }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg(const double* __restrict__ S, const double* __restrict__ rhs,
                                                      double* __restrict__ xout, int n, PcgWork wk, LMState* st,
                                                      unsigned int G) {
  __shared__ double p[kPcgMaxN];  // search direction (first n entries), identical in every workgroup
  __shared__ double Lf[kPcgMaxOwn][kBlk * kBlk];
  __shared__ double Sd[kPcgMaxOwn][kBlk * kBlk];
  __shared__ double dv[kPcgMaxOwn][kBlk];
  __shared__ int piv[kPcgMaxOwn][kBlk];
  __shared__ int kind[kPcgMaxOwn];
  __shared__ double xo[kPcgMaxOwn][kBlk], ro[kPcgMaxOwn][kBlk], zo[kPcgMaxOwn][kBlk], apo[kPcgMaxOwn][kBlk];
  __shared__ double slot;
  __shared__ int timed_out, singular;
  if (skip_solve(st)) return;
  const int min_iter = st->min_lin_iter, max_iter = st->max_lin_iter;
  const double tol = st->eta;
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const unsigned int g = blockIdx.x;
  const int nb = n / kBlk;
  const int nown = ((int)g < nb) ? (nb - (int)g + (int)G - 1) / (int)G : 0;
  double* part_rr = wk.part;
  double* part_rz = wk.part + kPcgMaxGroups;
  double* part_pap = wk.part + 2 * kPcgMaxGroups;
  unsigned int barriers = 0;
  int iterations = 0;
  if (tid == 0) {
    timed_out = 0;
    singular = 0;
  }
  // diagonal blocks of the owned rows, r = rhs, x = 0
  for (int i = tid; i < nown * kBlk * kBlk; i += kPcgThreads) {
    const int b = i / (kBlk * kBlk), e = i % (kBlk * kBlk);
    const int row = kBlk * (g + b * G) + e / kBlk, col = kBlk * (g + b * G) + e % kBlk;
    Lf[b][e] = S[(size_t)row * n + col];
    Sd[b][e] = Lf[b][e];
  }
  for (int i = tid; i < nown * kBlk; i += kPcgThreads) {
    const int b = i / kBlk, k = i % kBlk;
    ro[b][k] = rhs[kBlk * (g + b * G) + k];
    xo[b][k] = 0.0;
  }
  __syncthreads();
  if (tid < nown) {
    // Cholesky when the block is SPD, else an LU solve of the block
    if (chol9(Lf[tid], dv[tid])) {
      kind[tid] = 1;
    } else {
      for (int e = 0; e < kBlk * kBlk; ++e) Lf[tid][e] = Sd[tid][e];
      kind[tid] = lu9(Lf[tid], piv[tid], dv[tid]) ? 0 : -1;
      if (kind[tid] < 0) singular = 1;
    }
  }
  __syncthreads();
  // z = M^-1 r (own blocks), partials r.r and r.z
  if (tid < nown && kind[tid] >= 0) block_solve(Lf[tid], dv[tid], piv[tid], kind[tid], ro[tid], zo[tid]);
  __syncthreads();
  if (wave == 0) {
    double rr = 0.0, rz = 0.0;
    for (int i = lane; i < nown * kBlk; i += kWave) {
      const int b = i / kBlk, k = i % kBlk;
      const double r = ro[b][k], z = singular ? 0.0 : zo[b][k];
      rr += r * r;
      rz += r * z;
      wk.z[kBlk * (g + b * G) + k] = z;
    }
    rr = wave_sum(rr);
    rz = wave_sum(rz);
    if (lane == 0) {
      part_rr[g] = rr;
      part_rz[g] = rz;
    }
  }
  if (!grid_barrier(wk.bar, ++barriers * G, &timed_out)) goto fail;
  {
    const double rr0 = group_total(part_rr, G, &slot);
    double rz_old = group_total(part_rz, G, &slot);
    const double rhs_norm = fmax(sqrt(rr0), kEps);
    for (int c = tid; c < n; c += kPcgThreads) p[c] = wk.z[c];
    __syncthreads();
    if (!(min_iter <= 0 && sqrt(rr0) <= tol * rhs_norm)) {
      for (int k = 0; k < max_iter; ++k) {
        // Ap on the owned rows; wavefront w computes row w of each owned block
        for (int b = 0; b < nown; ++b) {
          const size_t row = (size_t)kBlk * (g + b * G) + wave;
          const double* Srow = S + row * n;
          double acc = 0.0;
          for (int c = lane; c < n; c += kWave) acc += Srow[c] * p[c];
          acc = wave_sum(acc);
          if (lane == 0) apo[b][wave] = acc;
        }
        __syncthreads();
        if (wave == 0) {
          double pap = 0.0;
          for (int i = lane; i < nown * kBlk; i += kWave) {
            const int b = i / kBlk, kk = i % kBlk;
            pap += p[kBlk * (g + b * G) + kk] * apo[b][kk];
          }
          pap = wave_sum(pap);
          if (lane == 0) part_pap[g] = pap;
        }
        if (!grid_barrier(wk.bar, ++barriers * G, &timed_out)) goto fail;
        const double denom = group_total(part_pap, G, &slot);
        if (fabs(denom) <= kEps) break;
        const double alpha = rz_old / denom;
        if (tid < nown * kBlk) {
          const int b = tid / kBlk, kk = tid % kBlk;
          xo[b][kk] += alpha * p[kBlk * (g + b * G) + kk];
          ro[b][kk] -= alpha * apo[b][kk];
        }
        iterations = k + 1;
        __syncthreads();
        if (tid < nown && kind[tid] >= 0) block_solve(Lf[tid], dv[tid], piv[tid], kind[tid], ro[tid], zo[tid]);
        __syncthreads();
        if (wave == 0) {
          double rr = 0.0, rz = 0.0;
          for (int i = lane; i < nown * kBlk; i += kWave) {
            const int b = i / kBlk, kk = i % kBlk;
            const double r = ro[b][kk], z = singular ? 0.0 : zo[b][kk];
            rr += r * r;
            rz += r * z;
            wk.z[kBlk * (g + b * G) + kk] = z;
          }
          rr = wave_sum(rr);
          rz = wave_sum(rz);
          if (lane == 0) {
            part_rr[g] = rr;
            part_rz[g] = rz;
          }
        }
        if (!grid_barrier(wk.bar, ++barriers * G, &timed_out)) goto fail;
        const double rr = group_total(part_rr, G, &slot);
        const double rz_new = group_total(part_rz, G, &slot);
        if (iterations >= min_iter && sqrt(rr) <= tol * rhs_norm) break;
        if (fabs(rz_old) <= kEps) break;
        const double beta = rz_new / rz_old;
        for (int c = tid; c < n; c += kPcgThreads) p[c] = wk.z[c] + beta * p[c];
        __syncthreads();
        rz_old = rz_new;
      }
    }
    for (int i = tid; i < nown * kBlk; i += kPcgThreads) {
      const int b = i / kBlk, kk = i % kBlk;
      xout[kBlk * (g + b * G) + kk] = xo[b][kk];
    }
    if (g == 0 && tid == 0) st->linear_iterations = iterations;
    if (tid == 0 && singular) st->step_invalid = 1;
    return;
  }
fail:
  if (tid == 0) st->step_invalid = 1;
}

}  // namespace

// solve scratch layout (doubles): [0,n) y, [n,2n) pivot reciprocals,
// [2n, 2n + nt kTB^2) diagonal-block inverses, then 2 nt + 2 ints of flags
size_t rcs_cholesky_tmp_doubles(int n) {
  const size_t nt = tri_blocks(n);
  return 2 * (size_t)n + nt * kTB * kTB + nt + 2;
}

namespace {
TriWork tri_work(double* tmp, int n) {
  const size_t nt = tri_blocks(n);
  TriWork tw;
  tw.y = tmp;
  tw.linv = tmp + 2 * (size_t)n;
  tw.flags = reinterpret_cast<int*>(tmp + 2 * (size_t)n + nt * kTB * kTB);
  tw.epoch = reinterpret_cast<unsigned int*>(tw.flags + 2 * nt);
  return tw;
}

int tri_groups(int n, int* groups) {
  int dev = 0, cus = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int nt = (int)tri_blocks(n);
  *groups = nt < cus ? nt : cus;
  if (*groups < 1) *groups = 1;
  return CT_OK;
}

int launch_tri_mt(const double* L, const double* rhs, double* x, int n, LMState* st, double* tmp, hipStream_t s) {
  int G = 0;
  CT_TRY(tri_groups(n, &G));
  hipLaunchKernelGGL(k_tri_solve_mt, dim3(G), dim3(kTriMtThreads), 0, s, L, rhs, x, n, tri_work(tmp, n), st, G);
  CT_LAUNCH_CHECK();
  return CT_OK;
}
}  // namespace

namespace {
__global__ void k_zero_words(double* p, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}
}  // namespace

int rcs_cholesky_prepare(double* tmp, int n, hipStream_t s) {
  // zero the flags and epoch of the many-workgroup triangular solves (a fill
  // kernel, so the call may sit inside a captured graph)
  const size_t nt = tri_blocks(n);
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, s, tmp + 2 * (size_t)n + nt * kTB * kTB, nt + 2);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_rcs_cholesky_solve(double* S, const double* rhs, double* x, int n, LMState* st, double* tmp,
                              hipStream_t s) {
  if (n <= 0) return CT_OK;
  if (n <= kSmallMax) {
    hipLaunchKernelGGL(k_chol_solve_small, dim3(1), dim3(kSmallThreads), 0, s, S, rhs, x, tmp, n, st);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
  for (int k0 = 0; k0 < n; k0 += kNB) {
    const int kb = n - k0 < kNB ? n - k0 : kNB;
    const int rows = n - k0 - kb;
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(kWave), 0, s, S, tmp + n, n, k0, st);
    CT_LAUNCH_CHECK();
    if (rows > 0) {
      const int grid = (rows + kPanelThreads - 1) / kPanelThreads;
      hipLaunchKernelGGL(k_chol_panel, dim3(grid), dim3(kPanelThreads), 0, s, S, tmp + n, n, k0, st);
      CT_LAUNCH_CHECK();
      const int T = (rows + kTile - 1) / kTile;
      hipLaunchKernelGGL(k_chol_syrk, dim3(T * (T + 1) / 2), dim3(256), 0, s, S, n, k0, kb, st);
      CT_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k_diag_inverse, dim3((unsigned int)tri_blocks(n)), dim3(kWave), 0, s, S,
                     tmp + 2 * (size_t)n, n, st);
  CT_LAUNCH_CHECK();
  if (rhs == nullptr) return CT_OK;  // factorisation only
  return launch_tri_mt(S, rhs, x, n, st, tmp, s);
}

int launch_rcs_tri_solve(const double* L, const double* rhs, double* x, int n, LMState* st, double* tmp,
                         hipStream_t s) {
  if (n <= 0) return CT_OK;
  if (n > kSmallMax) return launch_tri_mt(L, rhs, x, n, st, tmp, s);
  hipLaunchKernelGGL(k_tri_solve, dim3(1), dim3(kTriThreads), 0, s, L, rhs, x, tmp, n, st);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

size_t rcs_pcg_work_doubles(int n) { return (size_t)n + 3 * (size_t)kPcgMaxGroups + 8; }

unsigned int* rcs_pcg_barrier(double* work, int n) {
  return reinterpret_cast<unsigned int*>(work + n + 3 * (size_t)kPcgMaxGroups);
}

int rcs_pcg_groups(int n, int* groups) {
  const int nb = n / kBlk;
  int dev = 0, cus = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int G = nb < cus ? nb : cus;
  if (G > kPcgMaxGroups) G = kPcgMaxGroups;
  if (G < 1) G = 1;
  if ((nb + G - 1) / G > kPcgMaxOwn) {
    set_error("ITERATIVE_SCHUR supports at most " + std::to_string(kPcgMaxOwn * G) + " cameras on this device");
    return CT_ERR_UNSUPPORTED;
  }
  *groups = G;
  return CT_OK;
}

int launch_rcs_pcg(const double* S, const double* rhs, double* x, int n, LMState* st, double* work,
                   size_t work_doubles, hipStream_t s) {
  if (n <= 0) return CT_OK;
  if (n % kBlk != 0) {
    set_error("PCG expects 9-parameter camera blocks");
    return CT_ERR_ARGUMENT;
  }
  if (n > kPcgMaxN) {
    set_error("ITERATIVE_SCHUR supports at most " + std::to_string(kPcgMaxN / kBlk) + " cameras");
    return CT_ERR_UNSUPPORTED;
  }
  if (st == nullptr) {
    set_error("PCG needs the LM state (options and the cleared barrier)");
    return CT_ERR_ARGUMENT;
  }
  if (work_doubles < rcs_pcg_work_doubles(n)) {
    set_error("PCG workspace too small");
    return CT_ERR_WORKSPACE;
  }
  int G = 0;
  CT_TRY(rcs_pcg_groups(n, &G));
  PcgWork wk;
  wk.z = work;
  wk.part = work + n;
  wk.bar = rcs_pcg_barrier(work, n);
  hipLaunchKernelGGL(k_pcg, dim3(G), dim3(kPcgThreads), 0, s, S, rhs, x, n, wk, st, (unsigned int)G);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct
