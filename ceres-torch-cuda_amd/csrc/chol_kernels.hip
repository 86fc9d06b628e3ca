// FP64 solves of the reduced camera system S x = rhs (S symmetric positive
// definite, n x n row-major, lower triangle referenced).
//
//  * n <= 256: one workgroup factors S blockwise (16-column panels: the
//    diagonal block is factored in LDS by all threads, the rows below are
//    solved one row per thread and the trailing rank-16 update runs on 4x4
//    register tiles) and runs both triangular solves.
//  * n > 256: right-looking blocked Cholesky, 32-column panels; per panel one
//    workgroup factors the diagonal block, one kernel solves the rows below it
//    and one applies the trailing update on 64x64 tiles; a single workgroup
//    then runs the blocked triangular solves.
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
    y[i] = rhs[i];
  }
  __syncthreads();
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

// Factor the diagonal block of panel k0 in place (one workgroup) and store
// the pivots' reciprocals.
__global__ __launch_bounds__(kPanelThreads) void k_chol_diag(double* __restrict__ A, double* __restrict__ dinv_out,
                                                             int n, int k0, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double Ld[kNB][kNB + 1];
  __shared__ double dinv[kNB];
  __shared__ int bad;
  const int tid = threadIdx.x;
  const int kb = min(kNB, n - k0);
  if (tid == 0) bad = 0;
  for (int idx = tid; idx < kNB * kNB; idx += kPanelThreads) {
    const int r = idx / kNB, c = idx % kNB;
    Ld[r][c] = (r < kb && c <= r) ? A[(size_t)(k0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  factor_diag_block<kNB, kNB + 1, kPanelThreads>(Ld, dinv, kb, &bad);
  for (int idx = tid; idx < kb * kb; idx += kPanelThreads) {
    const int r = idx / kb, c = idx % kb;
    if (c <= r) A[(size_t)(k0 + r) * n + k0 + c] = Ld[r][c];
  }
  if (tid < kb) dinv_out[k0 + tid] = dinv[tid];
  if (tid == 0 && bad && st) st->step_invalid = 1;
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
// explicit-S PCG, one persistent kernel.  Workgroup g owns 9-row blocks
// g, g+G, ...  Reductions go through per-workgroup partials that every
// workgroup sums in the same order, so all of them take identical decisions.
// The grid barrier spins a bounded number of times; on timeout the solve is
// flagged invalid instead of waiting forever.
// ---------------------------------------------------------------------------
constexpr int kPcgThreads = 256;
constexpr int kBlk = 9;

struct PcgWork {
  double* x;        // [n]
  double* r;        // [n]
  double* z;        // [n]
  double* p;        // [n]
  double* Ap;       // [n]
  double* Lblk;     // [nb][81] Cholesky factors of diagonal blocks
  double* part;     // [G][4]
  unsigned int* bar;  // [2] count, generation
  int* flags;       // [nb] block is Cholesky-factored (1) or LU fallback (0)
};

__device__ bool grid_barrier(unsigned int* bar, unsigned int nblocks, int* timed_out) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned int gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int arrived = atomicAdd(&bar[0], 1u) + 1u;
    if (arrived == nblocks) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (++spins > (1L << 26)) {
          *timed_out = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();
  }
  __syncthreads();
  return *timed_out == 0;
}

__device__ void block_apply(const double* L, const double* rhs, double* out) {
  // solve (L L^T) out = rhs for a 9x9 Cholesky factor L (row-major lower)
  double t[kBlk];
#pragma unroll
  for (int i = 0; i < kBlk; ++i) {
    double s = rhs[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i * kBlk + k] * t[k];
    t[i] = s / L[i * kBlk + i];
  }
#pragma unroll
  for (int i = kBlk - 1; i >= 0; --i) {
    double s = t[i];
#pragma unroll
    for (int k = i + 1; k < kBlk; ++k) s -= L[k * kBlk + i] * out[k];
    out[i] = s / L[i * kBlk + i];
  }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg(const double* __restrict__ S, const double* __restrict__ rhs,
                                                      double* __restrict__ xout, int n, PcgWork wk, LMState* st,
                                                      int min_iter, int max_iter, double tol, unsigned int G) {
  if (skip_solve(st)) return;
  if (st != nullptr) {
    min_iter = st->min_lin_iter;
    max_iter = st->max_lin_iter;
    tol = st->eta;
  }
  __shared__ double red[kPcgThreads / kWave];
  __shared__ int timed_out;
  __shared__ double sums[4];
  __shared__ double L[81];
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const int nb = n / kBlk;
  if (tid == 0) timed_out = 0;
  __syncthreads();
  // factor own diagonal blocks; x = 0, r = rhs, z = M^-1 r, p = z
  double rr = 0.0, rz = 0.0, bb = 0.0;
  for (int blk = g; blk < nb; blk += G) {
    if (tid == 0) {
      bool ok = true;
      for (int i = 0; i < kBlk; ++i)
        for (int j = 0; j < kBlk; ++j) L[i * kBlk + j] = j <= i ? S[(size_t)(kBlk * blk + i) * n + kBlk * blk + j] : 0.0;
      for (int j = 0; j < kBlk && ok; ++j) {
        double d = L[j * kBlk + j];
        for (int k = 0; k < j; ++k) d -= L[j * kBlk + k] * L[j * kBlk + k];
        if (!(d > 0.0)) { ok = false; break; }
        d = sqrt(d);
        L[j * kBlk + j] = d;
        for (int i = j + 1; i < kBlk; ++i) {
          double s = L[i * kBlk + j];
          for (int k = 0; k < j; ++k) s -= L[i * kBlk + k] * L[j * kBlk + k];
          L[i * kBlk + j] = s / d;
        }
      }
      for (int k = 0; k < 81; ++k) wk.Lblk[81 * blk + k] = ok ? L[k] : 0.0;
      wk.flags[blk] = ok ? 1 : 0;
      double r[kBlk], zz[kBlk];
      for (int i = 0; i < kBlk; ++i) r[i] = rhs[kBlk * blk + i];
      if (ok) {
        block_apply(L, r, zz);
      } else {
        // non-SPD block: Jacobi on its diagonal (rare; the reference falls
        // back to an LU solve of the block)
        for (int i = 0; i < kBlk; ++i) {
          const double dgl = S[(size_t)(kBlk * blk + i) * n + kBlk * blk + i];
          zz[i] = r[i] / (fabs(dgl) > 0.0 ? dgl : 1.0);
        }
      }
      for (int i = 0; i < kBlk; ++i) {
        wk.x[kBlk * blk + i] = 0.0;
        wk.r[kBlk * blk + i] = r[i];
        wk.z[kBlk * blk + i] = zz[i];
        wk.p[kBlk * blk + i] = zz[i];
        rr += r[i] * r[i];
        rz += r[i] * zz[i];
        bb += r[i] * r[i];
      }
    }
  }
  if (tid == 0) {
    wk.part[4 * g] = rr;
    wk.part[4 * g + 1] = rz;
    wk.part[4 * g + 2] = bb;
  }
  if (!grid_barrier(wk.bar, G, &timed_out)) goto fail;
  {
    if (tid == 0) {
      double a = 0, b2 = 0, c = 0;
      for (unsigned int k = 0; k < G; ++k) {
        a += wk.part[4 * k];
        b2 += wk.part[4 * k + 1];
        c += wk.part[4 * k + 2];
      }
      sums[0] = a;
      sums[1] = b2;
      sums[2] = c;
    }
    __syncthreads();
    double rz_old = sums[1];
    const double rhs_norm = fmax(sqrt(sums[2]), 2.220446049250313e-16);
    double rnorm = sqrt(sums[0]);
    int iterations = 0;
    if (!(min_iter <= 0 && rnorm <= tol * rhs_norm)) {
      for (int k = 0; k < max_iter; ++k) {
        // Ap = S p on own rows, p.Ap partial
        double pap = 0.0;
        for (int blk = g; blk < nb; blk += G) {
          for (int rrow = 0; rrow < kBlk; ++rrow) {
            const int row = kBlk * blk + rrow;
            double s = 0.0;
            for (int c = tid; c < n; c += kPcgThreads) s += S[(size_t)row * n + c] * wk.p[c];
            const double tot = block_sum<kPcgThreads>(s, red);
            if (tid == 0) {
              wk.Ap[row] = tot;
              pap += wk.p[row] * tot;
            }
          }
        }
        if (tid == 0) wk.part[4 * g] = pap;
        if (!grid_barrier(wk.bar, G, &timed_out)) goto fail;
        if (tid == 0) {
          double a = 0;
          for (unsigned int q = 0; q < G; ++q) a += wk.part[4 * q];
          sums[3] = a;
        }
        __syncthreads();
        const double denom = sums[3];
        if (fabs(denom) <= 2.220446049250313e-16) break;
        const double alpha = rz_old / denom;
        // x, r update; z = M^-1 r; partial |r|^2, r.z
        double rr2 = 0.0, rz2 = 0.0;
        if (tid == 0) {
          for (int blk = g; blk < nb; blk += G) {
            double r[kBlk], zz[kBlk];
            for (int i = 0; i < kBlk; ++i) {
              const int row = kBlk * blk + i;
              wk.x[row] += alpha * wk.p[row];
              r[i] = wk.r[row] - alpha * wk.Ap[row];
              wk.r[row] = r[i];
              rr2 += r[i] * r[i];
            }
            if (wk.flags[blk]) {
              block_apply(wk.Lblk + 81 * blk, r, zz);
            } else {
              for (int i = 0; i < kBlk; ++i) {
                const double dgl = S[(size_t)(kBlk * blk + i) * n + kBlk * blk + i];
                zz[i] = r[i] / (fabs(dgl) > 0.0 ? dgl : 1.0);
              }
            }
            for (int i = 0; i < kBlk; ++i) {
              wk.z[kBlk * blk + i] = zz[i];
              rz2 += r[i] * zz[i];
            }
          }
          wk.part[4 * g + 1] = rr2;
          wk.part[4 * g + 2] = rz2;
        }
        iterations = k + 1;
        if (!grid_barrier(wk.bar, G, &timed_out)) goto fail;
        if (tid == 0) {
          double a = 0, b2 = 0;
          for (unsigned int q = 0; q < G; ++q) {
            a += wk.part[4 * q + 1];
            b2 += wk.part[4 * q + 2];
          }
          sums[0] = a;
          sums[1] = b2;
        }
        __syncthreads();
        rnorm = sqrt(sums[0]);
        if (iterations >= min_iter && rnorm <= tol * rhs_norm) break;
        const double rz_new = sums[1];
        if (fabs(rz_old) <= 2.220446049250313e-16) break;
        const double beta = rz_new / rz_old;
        if (tid == 0) {
          for (int blk = g; blk < nb; blk += G)
            for (int i = 0; i < kBlk; ++i) {
              const int row = kBlk * blk + i;
              wk.p[row] = wk.z[row] + beta * wk.p[row];
            }
        }
        rz_old = rz_new;
        if (!grid_barrier(wk.bar, G, &timed_out)) goto fail;
      }
    }
    for (int blk = g; blk < nb; blk += G)
      for (int i = tid; i < kBlk; i += kPcgThreads) {
        const double v = wk.x[kBlk * blk + i];
        xout[kBlk * blk + i] = v;
      }
    if (g == 0 && tid == 0 && st) st->linear_iterations = iterations;
    return;
  }
fail:
  if (tid == 0 && st) st->step_invalid = 1;
}

__global__ void k_reset_barrier(unsigned int* bar) {
  if (threadIdx.x < 2) bar[threadIdx.x] = 0u;
}

}  // namespace

size_t rcs_cholesky_tmp_doubles(int n) { return 2 * (size_t)n; }

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
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(kPanelThreads), 0, s, S, tmp + n, n, k0, st);
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
  hipLaunchKernelGGL(k_tri_solve, dim3(1), dim3(kTriThreads), 0, s, S, rhs, x, tmp, n, st);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_rcs_tri_solve(const double* L, const double* rhs, double* x, int n, LMState* st, double* tmp,
                         hipStream_t s) {
  if (n <= 0) return CT_OK;
  hipLaunchKernelGGL(k_tri_solve, dim3(1), dim3(kTriThreads), 0, s, L, rhs, x, tmp, n, st);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

size_t rcs_pcg_work_doubles(int n) {
  const int nb = n / kBlk;
  return (size_t)5 * n + (size_t)81 * nb + 4 * 1024 + 8 + nb;
}

int launch_rcs_pcg(const double* S, const double* rhs, double* x, int n, LMState* st, double* work,
                   size_t work_doubles, hipStream_t s) {
  if (n <= 0) return CT_OK;
  if (n % kBlk != 0) {
    set_error("PCG expects 9-parameter camera blocks");
    return CT_ERR_ARGUMENT;
  }
  if (work_doubles < rcs_pcg_work_doubles(n)) {
    set_error("PCG workspace too small");
    return CT_ERR_WORKSPACE;
  }
  const int nb = n / kBlk;
  int dev = 0, cus = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  unsigned int G = (unsigned int)(nb < cus ? nb : cus);
  if (G > 1024) G = 1024;
  PcgWork wk;
  wk.x = work;
  wk.r = work + n;
  wk.z = work + 2 * (size_t)n;
  wk.p = work + 3 * (size_t)n;
  wk.Ap = work + 4 * (size_t)n;
  wk.Lblk = work + 5 * (size_t)n;
  wk.part = wk.Lblk + 81 * (size_t)nb;
  wk.bar = reinterpret_cast<unsigned int*>(wk.part + 4 * 1024);
  wk.flags = reinterpret_cast<int*>(wk.part + 4 * 1024 + 8);
  hipLaunchKernelGGL(k_reset_barrier, dim3(1), dim3(64), 0, s, wk.bar);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pcg, dim3(G), dim3(kPcgThreads), 0, s, S, rhs, x, n, wk, st, 0, 500, 1e-10, G);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct
