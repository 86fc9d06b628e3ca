// FP64 solves of the reduced camera system S x = rhs (S symmetric positive
// definite, n x n row-major, lower triangle referenced).
//
//  * n <= 192: one workgroup of 8 waves keeps the matrix as 16x16 tiles in
//    matrix-core accumulator registers and the factor in LDS; the right-hand
//    side is factored along as an extra row (forward substitution), the
//    backward substitution runs from LDS.
//  * n > 192: left-looking Cholesky on 64x64 tiles in one persistent
//    launch (workgroups take tiles in dependency order and hand them on
//    through per-tile flags); the same launch forms V = L^-1 tile by tile, so
//    every solve with the factor is two bandwidth-bound products with V.
//  * explicit-S PCG with 9x9 block-Jacobi (SCHUR_JACOBI), diagonal (JACOBI) or
//    no (IDENTITY) preconditioning as one persistent kernel
//    (linear.py:1216-1256, preconditioners :462-537, 1156-1182).
#include <utility>

#include "ba_kernels.h"
#include "chol_small.h"
#include "pcg_device.h"

namespace ct {

namespace {

__device__ __forceinline__ bool skip_solve(const LMState* st) { return st != nullptr && (st->done || st->skip_step); }

// ---------------------------------------------------------------------------
// large n: one persistent launch for the factorisation and the inverse
//
// It works on kTB x kTB tiles (kTB = 64) and hands tiles between workgroups
// through per-tile flags stamped with the launch's epoch (read once at the
// start; the last workgroup to finish advances it), so flags never need
// clearing.  Workgroups take tiles from an atomic ticket counter in an order
// in which every tile depends only on tiles of smaller tickets.  A ticket is
// only taken by a running workgroup, so the oldest unfinished tile always has
// its inputs ready: every wait ends without any assumption about how many
// workgroups are resident at once.  Waits are bounded anyway; a timeout
// raises the launch's abort word, every workgroup stops, and the step is
// flagged invalid.
//
// Publication: every wave waits until its own stores have reached L2
// (s_waitcnt: a workgroup barrier alone does not wait for other waves'
// stores), then after the barrier one thread stamps the flag with an
// agent-scope release store, which writes L2 back for the other XCDs first.  A
// waiting workgroup's polling thread uses plain atomic loads and one
// agent-scope acquire once the flag is seen; the caches it invalidates (L1
// per CU, L2 per XCD) are the ones every wave of the workgroup reads through.
// ---------------------------------------------------------------------------
constexpr int kTB = 64;
constexpr int kFThreads = 256;
constexpr int kFLd = kTB + 2;  // LDS row stride: 16-byte aligned rows, 16 rows on distinct banks
constexpr long kSpinLimit = 1L << 24;

__host__ __device__ inline size_t tri_blocks(int n) { return ((size_t)n + kTB - 1) / kTB; }
__host__ __device__ inline size_t factor_tiles(int n) {
  const size_t nt = tri_blocks(n);
  return nt * (nt + 1) / 2;
}

// x = S^-1 rhs (rhs == nullptr: factor only); the factor overwrites the lower
// triangle of A and the pivots' reciprocals go to tmp[n:2n].
__global__ __launch_bounds__(kSmallThreads) void k_chol_small(double* __restrict__ A, const double* __restrict__ rhs,
                                                              double* __restrict__ x, double* __restrict__ tmp, int n,
                                                              LMState* st) {
  if (skip_solve(st)) return;
  __shared__ SmallCholShared sh;
  const int tid = threadIdx.x;
  small_factor(A, rhs, n, sh);
  small_store_lower(A, sh.Lp, n);
  for (int i = tid; i < n; i += kSmallThreads) tmp[n + i] = sh.dinv[i];
  if (rhs == nullptr) {
    if (tid == 0 && st && sh.bad) st->step_invalid = 1;
    return;
  }
  double* y = sh.Lp + packed_index(n, 0);  // row n: y = L^-1 rhs
  small_backward(sh.Lp, sh.dinv, y, n);
  int my_bad = 0;
  for (int i = tid; i < n; i += kSmallThreads) {
    x[i] = y[i];
    if (!isfinite(y[i])) my_bad = 1;
  }
  if (my_bad) sh.bad = 1;
  __syncthreads();
  if (tid == 0 && st) {
    if (sh.bad) st->step_invalid = 1;
    st->linear_iterations = 1;
  }
}

// Both triangular solves with a factor left by k_chol_small (lower triangle
// of L, pivots' reciprocals in tmp[n:2n]): the factor is loaded into LDS once.
__global__ __launch_bounds__(kSmallThreads) void k_tri_small(const double* __restrict__ L,
                                                             const double* __restrict__ rhs, double* __restrict__ x,
                                                             const double* __restrict__ tmp, int n, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double Lp[kSmallPacked];
  __shared__ double dinv[kSmallMax];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  small_load_lower(Lp, L, n);
  double* y = Lp + packed_index(n, 0);
  for (int i = tid; i < n; i += kSmallThreads) {
    dinv[i] = tmp[n + i];
    y[i] = rhs[i];
  }
  __syncthreads();
  small_forward(Lp, dinv, y, n);
  small_backward(Lp, dinv, y, n);
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

// Every thread calls this; thread 0 polls up to two flags.  *status (LDS) is
// written by thread 0 before the barrier and read by all threads after it;
// callers keep a barrier between two calls.
__device__ __forceinline__ bool wait_flags(const int* f1, const int* f2, int epoch, unsigned int* abort_word,
                                           int* status) {
  if (threadIdx.x == 0 && *status == 0) {
    long spins = 0;
    for (const int* f : {f1, f2}) {
      if (f == nullptr) continue;
      while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (__hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (unsigned int)epoch ||
            ++spins > kSpinLimit) {
          *status = 1;
          __hip_atomic_store(abort_word, (unsigned int)epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (*status) break;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return *status == 0;
}

__device__ __forceinline__ void publish_flag(int* flag, int epoch) {
  own_stores_done();
  __syncthreads();
  // the release store writes L2 back (for the other XCDs) before the flag
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Next ticket for this workgroup (every thread calls it), or `stop` once the
// launch is aborted.  The barrier in front keeps the previous tile's LDS use
// and the previous ticket's readers ahead of the overwrite.
__device__ __forceinline__ int take_ticket(unsigned int* counter, int stop, const int* status, int* slot) {
  __syncthreads();
  if (threadIdx.x == 0) *slot = *status ? stop : (int)atomicAdd(counter, 1u);
  __syncthreads();
  return *slot;
}

// The last workgroup of the launch to finish resets the ticket counter and
// advances the epoch for the next launch on the stream.
__device__ __forceinline__ void finish_launch(unsigned int* ctl, int epoch) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned int done = atomicAdd(&ctl[1], 1u);
    if (done == gridDim.x - 1) {
      ctl[1] = 0u;
      ctl[2] = 0u;
      __threadfence();
      __hip_atomic_store(&ctl[0], (unsigned int)epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Copy a kTB x kTB tile (rows r0.., columns c0.. of a row-major matrix with
// leading dimension ld; rows >= rows_valid read as zero) into LDS.  All 16
// loads of a thread are issued before the first store.
__device__ __forceinline__ void stage_tile(double (*dst)[kFLd], const double* __restrict__ src, int ld, int r0,
                                           int c0, int rows_valid, int tid) {
  constexpr int kPer = kTB * kTB / kFThreads;
  double v[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int r = u * (kFThreads / kTB) + tid / kTB, c = tid % kTB;
    v[u] = r < rows_valid ? src[(size_t)(r0 + r) * ld + c0 + c] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) dst[u * (kFThreads / kTB) + tid / kTB][tid % kTB] = v[u];
}

// One doubling step of the blocked inverse of the lower-triangular factor in
// D (V holds the inverses of the diagonal blocks of width H, zero above the
// diagonal): for every 2H-block at base b, X21 = -V22 (L21 V11).  The
// products run on the matrix cores as 16x16 output blocks, one per wave,
// over full block widths (the zero upper triangles make the triangular
// bounds implicit), with T = L21 V11 staged in D's unused upper triangle at
// D[b + r][b + H + c].
template <int H>
__device__ __forceinline__ void inverse_doubling(double (*D)[kFLd], double (*V)[kFLd], int w, int lane) {
  constexpr int kSub = H / 16, kTasks = (kTB / (2 * H)) * kSub * kSub;
  const int li = lane & 15, lk = lane >> 4;
  for (int t = w; t < kTasks; t += kFThreads / kWave) {
    const int b = (t / (kSub * kSub)) * 2 * H, r0 = 16 * ((t / kSub) % kSub), c0 = 16 * (t % kSub);
    dx4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < H; k0 += 4)
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(D[b + H + r0 + li][b + k0 + lk], V[b + k0 + lk][b + c0 + li], c, 0,
                                               0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) D[b + r0 + 4 * v + lk][b + H + c0 + li] = c[v];
  }
  __syncthreads();
  for (int t = w; t < kTasks; t += kFThreads / kWave) {
    const int b = (t / (kSub * kSub)) * 2 * H, r0 = 16 * ((t / kSub) % kSub), c0 = 16 * (t % kSub);
    dx4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < H; k0 += 4)
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(-V[b + H + r0 + li][b + H + k0 + lk], D[b + k0 + lk][b + H + c0 + li],
                                               c, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) V[b + H + r0 + 4 * v + lk][b + c0 + li] = c[v];
  }
  __syncthreads();
}

struct FactorWork {
  double* dinv;        // [n] reciprocals of the pivots
  double* linv;        // [nt][kTB][kTB] inverses of the diagonal tiles of L (row-major, zero outside)
  double* vinv;        // [n][n] V = L^-1: tiles below the diagonal (row-major, leading dimension n)
  int* ready;          // [nt (nt + 1) / 2] epoch stamp of each finished tile of L
  int* vready;         // [nt (nt - 1) / 2] epoch stamp of each finished tile of V below the diagonal
  unsigned int* ctl;   // [0] completed launches, [1] finished workgroups, [2] next ticket, [3] abort epoch
};

// Lower tiles in column-major order: column j holds tiles (j..nt-1, j).
__device__ __forceinline__ int tile_index(int i, int j, int nt) { return j * nt - j * (j - 1) / 2 + (i - j); }
// Tiles strictly below the diagonal in column-major order: column j holds (j+1..nt-1, j).
__device__ __forceinline__ int vtile_index(int i, int j, int nt) { return j * (nt - 1) - j * (j - 1) / 2 + (i - j - 1); }

// 64x64 tile products on the matrix cores.  Wave w of the workgroup owns the
// 32x32 quadrant (w / 2, w % 2) as 2x2 blocks of v_mfma_f64_16x16x4f64: in
// lane l, block b = 2 bi + bj holds in item v the element
//   (32 (w / 2) + 16 bi + 4 v + l / 16,  32 (w % 2) + 16 bj + l % 16).
// The A operand of one instruction is a 16x4 slice (lane l: row l % 16,
// column l / 16), the B operand a 4x16 slice (lane l: row l / 16, column
// l % 16).
__device__ __forceinline__ int acc_row(int w, int b, int v, int lane) {
  return 32 * (w >> 1) + 16 * (b >> 1) + 4 * v + (lane >> 4);
}
__device__ __forceinline__ int acc_col(int w, int b, int lane) { return 32 * (w & 1) + 16 * (b & 1) + (lane & 15); }

__device__ __forceinline__ void acc_zero(dx4 (&acc)[4]) {
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = dx4{0.0, 0.0, 0.0, 0.0};
}

// acc += s P Q over k = 0..kTB-1, P[r][k] = Pm[r][k] and Q[k][c] = Qm[c][k]
// (kQT) or Qm[k][c]; Pm and Qm are kTB x kTB LDS tiles, s is +1 or -1.
template <bool kQT>
__device__ __forceinline__ void tile_mma(dx4 (&acc)[4], const double (*Pm)[kFLd], const double (*Qm)[kFLd], double s,
                                         int w, int lane) {
  const int li = lane & 15, lk = lane >> 4;
  const int pr = 32 * (w >> 1) + li, qc = 32 * (w & 1) + li;
#pragma unroll 4
  for (int k0 = 0; k0 < kTB; k0 += 4) {
    const int k = k0 + lk;
    const double a0 = s * Pm[pr][k], a1 = s * Pm[pr + 16][k];
    const double b0 = kQT ? Qm[qc][k] : Qm[k][qc];
    const double b1 = kQT ? Qm[qc + 16][k] : Qm[k][qc + 16];
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[3], 0, 0, 0);
  }
}

// Rank-16 update of the diagonal tile's trailing lower triangle after the
// panel at column `base`: D[r][c] -= sum_q D[r][base+q] D[c][base+q] for
// base+16 <= c <= r, as 16x16 blocks on the matrix cores (block (bi, bj),
// bi >= bj, of the trailing part goes to wave (block index mod 4)).
__device__ __forceinline__ void panel_trailing_update(double (*D)[kFLd], int base, int w, int lane) {
  const int t0 = base + kPW, tb = (kTB - t0) / kPW;  // trailing 16-blocks per side
  const int li = lane & 15, lk = lane >> 4;
  for (int blk = w; blk < tb * (tb + 1) / 2; blk += kFThreads / kWave) {
    int bi = 0;
    while ((bi + 1) * (bi + 2) / 2 <= blk) ++bi;
    const int bj = blk - bi * (bi + 1) / 2;
    const int r0 = t0 + kPW * bi, c0 = t0 + kPW * bj;
    dx4 c;
#pragma unroll
    for (int v = 0; v < 4; ++v) c[v] = D[r0 + 4 * v + lk][c0 + li];
#pragma unroll
    for (int k0 = 0; k0 < kPW; k0 += 4) {
      const double a = -D[r0 + li][base + k0 + lk];
      const double b = D[c0 + li][base + k0 + lk];
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) D[r0 + 4 * v + lk][c0 + li] = c[v];
  }
}

// Tile (i, j), i > j, of V = L^-1: acc = sum_{k=j}^{i-1} L_ik V_kj, then
// V_ij = -V_ii acc.  V_jj and V_ii are the diagonal tiles' inverses.  tid is
// the caller's thread index.  Returns false once the launch is aborted.  Every
// thread calls it.
__device__ __forceinline__ bool inverse_tile(const double* __restrict__ A, int n, int nt, int i, int j, int tid, const FactorWork& fw,
                             int epoch, double (*lds)[kTB][kFLd], int* status) {
  const int w = tid / kWave, lane = tid % kWave;
  const int r0 = i * kTB, ri = min(kTB, n - r0);
  dx4 acc[4];
  acc_zero(acc);
  double(*Li)[kFLd] = lds[0];
  double(*Vk)[kFLd] = lds[1];
  for (int k = j; k < i; ++k) {
    if (!wait_flags(&fw.ready[tile_index(i, k, nt)],
                    k == j ? &fw.ready[tile_index(j, j, nt)] : &fw.vready[vtile_index(k, j, nt)], epoch, &fw.ctl[3],
                    status))
      return false;
    stage_tile(Li, A, n, r0, k * kTB, ri, tid);  // L_ik (k < i: full width)
    if (k == j) stage_tile(Vk, fw.linv + (size_t)j * kTB * kTB, kTB, 0, 0, kTB, tid);
    else stage_tile(Vk, fw.vinv, n, k * kTB, j * kTB, kTB, tid);  // V_kj, j < k < i: full tile
    __syncthreads();
    tile_mma<false>(acc, Li, Vk, 1.0, w, lane);
    __syncthreads();
  }
  if (!wait_flags(&fw.ready[tile_index(i, i, nt)], nullptr, epoch, &fw.ctl[3], status)) return false;
  double(*X)[kFLd] = lds[0];
  double(*Vi)[kFLd] = lds[1];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int v = 0; v < 4; ++v) X[acc_row(w, b, v, lane)][acc_col(w, b, lane)] = acc[b][v];
  stage_tile(Vi, fw.linv + (size_t)i * kTB * kTB, kTB, 0, 0, kTB, tid);
  __syncthreads();
  dx4 out[4];
  acc_zero(out);
  tile_mma<false>(out, Vi, X, -1.0, w, lane);
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = acc_row(w, b, v, lane), col = acc_col(w, b, lane);
      if (row < ri) fw.vinv[(size_t)(r0 + row) * n + j * kTB + col] = out[b][v];
    }
  publish_flag(&fw.vready[vtile_index(i, j, nt)], epoch);
  return true;
}

// Left-looking tiled Cholesky, S = L L^T in place (lower triangle), in one
// launch.  The workgroup holding tile (i, j) keeps it in registers (matrix-
// core accumulators) and subtracts L_ik L_jk^T for k = 0..j-1 as those tiles
// are published, always in increasing k, so the arithmetic does not depend on
// which workgroup takes which tile.  Then:
//  * diagonal tile: wave 0 factors it in registers (lane l owns row l); the
//    inverse of the factor is formed blockwise (16x16 diagonal blocks by
//    substitution, then two doubling steps X21 = -V22 (L21 V11)) and kept for
//    the off-diagonal tiles of the column and for the triangular solves; the
//    tile is published once its inverse is stored (no other tile reads the
//    diagonal tile of L) and its factor entries are stored after that;
//  * off-diagonal tile: L_ij = A_ij V_jj^T with the published inverse.
// Rows past n in the last tile row are zero; the diagonal tile pads them with
// an identity block.  A pivot that is not positive and finite flags the step.
// Row c of V = L^-1 below the diagonal follows column c of L in the ticket
// order (workgroups that would otherwise idle on the factor's critical path
// take those tiles): V_ij = -V_ii sum_{k=j}^{i-1} L_ik V_kj, k in increasing
// order.  Each later solve is then two products with V.
__global__ __launch_bounds__(kFThreads) void k_chol_tiles(double* __restrict__ A, int n, FactorWork fw,
                                                          LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double lds[2][kTB][kFLd];
  __shared__ double pinv[kTB];
  __shared__ int s_status, s_ticket, s_bad;
  const int nt = (int)tri_blocks(n);
  const int nall = nt * nt;
  const int epoch = (int)__hip_atomic_load(&fw.ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
  if (threadIdx.x == 0) s_status = 0;
  for (;;) {
    // tickets c nt .. c nt + nt - 1: column c of L (nt - c tiles), then row c
    // of V below the diagonal (c tiles); everything a tile reads comes from
    // a smaller ticket
    const int q = take_ticket(&fw.ctl[2], nall, &s_status, &s_ticket);
    if (q >= nall) break;
    const int seg = q / nt, pos = q % nt;
    if (pos >= nt - seg) {
      if (!inverse_tile(A, n, nt, seg, pos - (nt - seg), (int)threadIdx.x ^ (q >> 30), fw, epoch, lds, &s_status))
        break;
      continue;
    }
    // a copy of the thread index the compiler cannot see through (q < 2^30,
    // so q >> 30 is zero): nothing derived from it (LDS and global addresses)
    // is computed ahead and kept live across tiles
    const int tid = (int)threadIdx.x ^ (q >> 30);
    const int w = tid / kWave, lane = tid % kWave;
    const int i = seg + pos, j = seg;
    const int r0 = i * kTB, c0 = j * kTB;
    const int ri = min(kTB, n - r0), cj = min(kTB, n - c0);
    dx4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = acc_row(w, b, v, lane), col = acc_col(w, b, lane);
        acc[b][v] = (row < ri && col < cj) ? A[(size_t)(r0 + row) * n + c0 + col] : 0.0;
      }
    double(*Li)[kFLd] = lds[0];
    double(*Lj)[kFLd] = (i == j) ? lds[0] : lds[1];
    for (int k = 0; k < j; ++k) {
      if (!wait_flags(&fw.ready[tile_index(i, k, nt)], i == j ? nullptr : &fw.ready[tile_index(j, k, nt)], epoch,
                      &fw.ctl[3], &s_status))
        break;
      const int kk = k * kTB;  // tiles left of the diagonal are full width
      stage_tile(Li, A, n, r0, kk, ri, tid);
      if (i != j) stage_tile(Lj, A, n, c0, kk, kTB, tid);
      __syncthreads();
      tile_mma<true>(acc, Li, Lj, -1.0, w, lane);
      __syncthreads();
    }
    if (s_status) break;
    if (i == j) {
      double(*D)[kFLd] = lds[0];
      double(*V)[kFLd] = lds[1];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = acc_row(w, b, v, lane), col = acc_col(w, b, lane);
          D[row][col] = col > row ? 0.0 : (row < cj ? acc[b][v] : (row == col ? 1.0 : 0.0));
          V[row][col] = 0.0;
        }
      __syncthreads();
      // right-looking over four 16-column panels: wave 0 factors the panel
      // (rows base..63) in registers, then all threads apply its rank-16
      // update to the trailing lower triangle in LDS
      if (tid == 0) s_bad = 0;
      for (int base = 0; base < kTB; base += kPW) {
        if (tid < kWave) {
          const int l = tid, row = base + l;
          double a[kPW];
#pragma unroll
          for (int q = 0; q < kPW; ++q) a[q] = row < kTB ? D[row][base + q] : 0.0;
          lds_wave_sync();
          double pin = 0.0;
          int pbad = 0;
          panel_steps_ahead<true>(a, pin, pbad, l, kPW, std::make_integer_sequence<int, kPW>{});
          if (l < kPW) pinv[base + l] = pin;
          if (pbad && l == 0) s_bad = 1;
          if (row < kTB) {
#pragma unroll
            for (int q = 0; q < kPW; ++q) D[row][base + q] = a[q];
          }
        }
        __syncthreads();
        panel_trailing_update(D, base, w, lane);
        __syncthreads();
      }
      if (tid < cj) fw.dinv[c0 + tid] = pinv[tid];
      if (tid == 0 && s_bad && st) st->step_invalid = 1;
      // inverse, 16x16 diagonal blocks: wave w inverts block w, lane c < 16 one column
      {
        const int w = tid / kWave, c = tid % kWave, b0 = 16 * w;
        if (c < 16) {
          double x[16];
#pragma unroll
          for (int ii = 0; ii < 16; ++ii) {
            double s = ii == c ? 1.0 : 0.0;
#pragma unroll
            for (int kq = 0; kq < ii; ++kq) s = fma(-D[b0 + ii][b0 + kq], x[kq], s);
            x[ii] = ii < c ? 0.0 : s * pinv[b0 + ii];
          }
#pragma unroll
          for (int ii = 0; ii < 16; ++ii) V[b0 + ii][b0 + c] = x[ii];
        }
      }
      __syncthreads();
      // doubling steps: for blocks of width h at base b, X21 = -V22 (L21 V11);
      // T = L21 V11 is staged in the (unused) upper triangle of D at D[b + r][b + h + c]
      inverse_doubling<16>(D, V, w, lane);
      inverse_doubling<32>(D, V, w, lane);
      double* vout = fw.linv + (size_t)j * kTB * kTB;
      for (int idx = tid; idx < kTB * kTB; idx += kFThreads) {
        const int r = idx / kTB, c = idx % kTB;
        vout[idx] = (r < cj && c < cj) ? V[r][c] : 0.0;
      }
      // other tiles read the inverse and the pivots, never this tile of L:
      // publish first, then store the factor (complete when the launch ends)
      publish_flag(&fw.ready[tile_index(j, j, nt)], epoch);
      for (int idx = tid; idx < kTB * kTB; idx += kFThreads) {
        const int r = idx / kTB, c = idx % kTB;
        if (r < cj && c <= r) A[(size_t)(c0 + r) * n + c0 + c] = D[r][c];
      }
      continue;
    } else {
      if (!wait_flags(&fw.ready[tile_index(j, j, nt)], nullptr, epoch, &fw.ctl[3], &s_status)) break;
      double(*X)[kFLd] = lds[0];
      double(*Vt)[kFLd] = lds[1];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) X[acc_row(w, b, v, lane)][acc_col(w, b, lane)] = acc[b][v];
      stage_tile(Vt, fw.linv + (size_t)j * kTB * kTB, kTB, 0, 0, kTB, tid);
      __syncthreads();
      dx4 out[4];
      acc_zero(out);
      tile_mma<true>(out, X, Vt, 1.0, w, lane);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = acc_row(w, b, v, lane), col = acc_col(w, b, lane);
          if (row < ri) A[(size_t)(r0 + row) * n + c0 + col] = out[b][v];
        }
    }
    publish_flag(&fw.ready[tile_index(i, j, nt)], epoch);
  }
  if (threadIdx.x == 0 && s_status && st) st->step_invalid = 1;
  finish_launch(fw.ctl, epoch);
}

// Solves with the inverse: y = V b, x = V^T y.  Each lower tile (i, j) of V
// (the diagonal tiles from linv) writes its partial product to
// part[block][row]; a second kernel sums the partials of each row in one
// fixed order, so the solve is deterministic and reads V once per product.
constexpr int kApplyThreads = 256;

// part[j][i kTB + r] = sum_c V_ij[r][c] vec[j kTB + c]            (transpose = false)
// part[i][j kTB + c] = sum_r V_ij[r][c] vec[i kTB + r]            (transpose = true)
template <bool kTranspose>
__global__ __launch_bounds__(kApplyThreads) void k_inv_partials(const double* __restrict__ vinv,
                                                                 const double* __restrict__ linv,
                                                                 const double* __restrict__ vec,
                                                                 double* __restrict__ part, int n, LMState* st) {
  if (skip_solve(st)) return;
  __shared__ double red[kApplyThreads / kTB][kTB];
  const int nt = (int)tri_blocks(n);
  const int q = blockIdx.x;
  int j = 0, rest = q;
  while (rest >= nt - j) {
    rest -= nt - j;
    ++j;
  }
  const int i = j + rest;
  const int tid = threadIdx.x;
  const int r0 = i * kTB, c0 = j * kTB;
  const int ri = min(kTB, n - r0), cj = min(kTB, n - c0);
  const double* tile = (i == j) ? linv + (size_t)j * kTB * kTB : vinv + (size_t)r0 * n + c0;
  const int ld = (i == j) ? kTB : n;
  if (!kTranspose) {
    // thread (row r, 16-column slice s): contiguous 128-byte reads
    const int r = tid / 4, sl = tid % 4;
    double sum = 0.0;
    if (r < ri) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int c = sl * 16 + u;
        if (c < cj) sum = fma(tile[(size_t)r * ld + c], vec[c0 + c], sum);
      }
    }
    sum += __shfl_xor(sum, 1, kWave);
    sum += __shfl_xor(sum, 2, kWave);
    if (sl == 0 && r < ri) part[(size_t)j * n + r0 + r] = sum;
  } else {
    // thread (16-row slice s, column c): coalesced along c
    const int sl = tid / kTB, c = tid % kTB;
    double sum = 0.0;
    if (c < cj) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int r = sl * 16 + u;
        if (r < ri) sum = fma(tile[(size_t)r * ld + c], vec[r0 + r], sum);
      }
    }
    red[sl][c] = sum;
    __syncthreads();
    if (tid < cj) {
      double t = 0.0;
#pragma unroll
      for (int qq = 0; qq < kApplyThreads / kTB; ++qq) t += red[qq][tid];
      part[(size_t)i * n + c0 + tid] = t;
    }
  }
}

// out[e] = sum over the partial blocks of e in increasing block order: blocks
// 0..b(e) for V b, blocks b(e)..nt-1 for V^T y (b(e) = e / kTB).
template <bool kTranspose>
__global__ void k_inv_reduce(const double* __restrict__ part, double* __restrict__ out, int n, LMState* st) {
  if (skip_solve(st)) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int nt = (int)tri_blocks(n), be = e / kTB;
  double t = 0.0;
  if (!kTranspose) {
    for (int b = 0; b <= be; ++b) t += part[(size_t)b * n + e];
  } else {
    for (int b = be; b < nt; ++b) t += part[(size_t)b * n + e];
    if (!isfinite(t) && st) st->step_invalid = 1;
  }
  out[e] = t;
  if (kTranspose && e == 0 && st) st->linear_iterations = 1;
}

// ---------------------------------------------------------------------------
// explicit-S PCG, one persistent kernel (linear.py:1216-1256); preconditioner
// SCHUR_JACOBI (linear.py:1156-1182, applied as in linear.py:69-77), JACOBI
// (the inverse diagonal, linear.py:531-537) or IDENTITY (linear.py:54-55).  Workgroup g owns the 9-row camera blocks g, g+G, ...;
// its 9 wavefronts compute one row each of S p.  Every workgroup keeps the
// whole search direction p in LDS and updates it identically from the
// shared z, so an iteration needs two grid barriers (after p.Ap and after
// r.z).  Cross-workgroup sums read the per-workgroup partials in one fixed
// order in every workgroup, so all of them take identical decisions.  The
// grid barrier is a flag array: each workgroup stores the barrier's number
// (the solve's epoch in the high word, the barrier index in the low word) in
// its own slot and one wave polls all slots, so no two workgroups contend for
// one address (a single arrival counter serialised 200 atomics, 13-18 us per
// barrier).  The LM begin kernel advances the epoch before each solve; waits
// are bounded and a timeout flags the step invalid.
// ---------------------------------------------------------------------------
constexpr int kPcgWaves = kBlk;  // one wavefront per row of a camera block
constexpr int kPcgThreads = kPcgWaves * kWave;
constexpr int kPcgMaxOwn = 8;    // camera blocks per workgroup
constexpr int kPcgMaxGroups = 1024;
constexpr int kPcgMaxN = 12288;  // p is held in LDS

struct PcgWork {
  double* z;                  // [n]  preconditioned residual, read by every workgroup
  double* part;               // [3][kPcgMaxGroups] partials: r.r, r.z, p.Ap
  unsigned int* epoch;        // solves so far (advanced by the LM begin kernel)
  unsigned long long* flags;  // [kPcgMaxGroups] last barrier passed by each workgroup
};

// One row of S p: lane l sums S[row][c] p[c] over c = l, l + kWave, ... in
// that order, with the loads of kRowBatch terms issued before their adds.
constexpr int kRowBatch = 16;
__device__ __forceinline__ double row_dot(const double* __restrict__ Srow, const double* p, int n, int lane) {
  double acc = 0.0;
  for (int c0 = lane; c0 < n; c0 += kRowBatch * kWave) {
    double sv[kRowBatch];
#pragma unroll
    for (int t = 0; t < kRowBatch; ++t) {
      const int c = c0 + t * kWave;
      sv[t] = c < n ? Srow[c] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < kRowBatch; ++t) {
      const int c = c0 + t * kWave;
      if (c < n) acc += sv[t] * p[c];
    }
  }
  return acc;
}

// p = z + beta p over the whole vector (every workgroup), loads issued first.
constexpr int kVecBatch = 8;
// The first kVecBatch entries of z a thread updates, loaded ahead (issued
// before the group totals, so both wait out one latency together).
__device__ __forceinline__ void load_direction_head(double (&zv)[kVecBatch], const double* __restrict__ z, int n) {
#pragma unroll
  for (int t = 0; t < kVecBatch; ++t) {
    const int c = threadIdx.x + t * kPcgThreads;
    zv[t] = c < n ? z[c] : 0.0;
  }
}

// p = z (copy) or p = z + beta p, entries in the same order as a plain pass;
// zv holds the head from load_direction_head.
__device__ __forceinline__ void update_direction(double* p, const double (&zv)[kVecBatch],
                                                 const double* __restrict__ z, double beta, int n, bool copy) {
#pragma unroll
  for (int t = 0; t < kVecBatch; ++t) {
    const int c = threadIdx.x + t * kPcgThreads;
    if (c < n) p[c] = copy ? zv[t] : zv[t] + beta * p[c];
  }
  for (int c0 = threadIdx.x + kVecBatch * kPcgThreads; c0 < n; c0 += kVecBatch * kPcgThreads) {
    double zv[kVecBatch];
#pragma unroll
    for (int t = 0; t < kVecBatch; ++t) {
      const int c = c0 + t * kPcgThreads;
      zv[t] = c < n ? z[c] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < kVecBatch; ++t) {
      const int c = c0 + t * kPcgThreads;
      if (c < n) p[c] = copy ? zv[t] : zv[t] + beta * p[c];
    }
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
  __shared__ double slot, slots2[2];
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
  // barrier numbers of this solve: epoch in the high word, index in the low word
  const unsigned long long bar_base =
      (unsigned long long)__hip_atomic_load(wk.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) << 32;
  unsigned int barriers = 0;
  int iterations = 0;
  if (tid == 0) {
    timed_out = 0;
    singular = 0;
  }
  // diagonal blocks of the owned rows, r = rhs, x = 0; with split blocks
  // (quaternion cameras: the rotation and the other six values are separate
  // parameter blocks) the rotation / rest coupling is left out of the block
  const bool split = st->split_blocks != 0;
  for (int i = tid; i < nown * kBlk * kBlk; i += kPcgThreads) {
    const int b = i / (kBlk * kBlk), e = i % (kBlk * kBlk);
    const int row = kBlk * (g + b * G) + e / kBlk, col = kBlk * (g + b * G) + e % kBlk;
    const bool cross = split && ((e / kBlk < 3) != (e % kBlk < 3));
    Lf[b][e] = cross ? 0.0 : S[(size_t)row * n + col];
    Sd[b][e] = Lf[b][e];
  }
  for (int i = tid; i < nown * kBlk; i += kPcgThreads) {
    const int b = i / kBlk, k = i % kBlk;
    ro[b][k] = rhs[kBlk * (g + b * G) + k];
    xo[b][k] = 0.0;
  }
  __syncthreads();
  if (tid < nown) {
    if (st->precond == CT_PRECONDITIONER_JACOBI) {
      // 1 / max(S_kk, eps) (linear.py:531-537)
      for (int k = 0; k < kBlk; ++k) dv[tid][k] = 1.0 / fmax(Sd[tid][k * kBlk + k], kEps);
      kind[tid] = 2;
    } else if (st->precond == CT_PRECONDITIONER_IDENTITY) {
      kind[tid] = 3;
    } else if (chol9(Lf[tid], dv[tid])) {
      // Cholesky when the block is SPD, else an LU solve of the block
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
  if (!grid_barrier(wk.flags, bar_base + ++barriers, G, &timed_out)) goto fail;
  {
    double zv[kVecBatch];
    load_direction_head(zv, wk.z, n);
    double rr0, rz_old;
    group_totals2(part_rr, part_rz, G, slots2, &rr0, &rz_old);
    const double rhs_norm = fmax(sqrt(rr0), kEps);
    update_direction(p, zv, wk.z, 0.0, n, true);
    __syncthreads();
    if (!(min_iter <= 0 && sqrt(rr0) <= tol * rhs_norm)) {
      for (int k = 0; k < max_iter; ++k) {
        // Ap on the owned rows; wavefront w computes row w of each owned block
        for (int b = 0; b < nown; ++b) {
          const size_t row = (size_t)kBlk * (g + b * G) + wave;
          double acc = row_dot(S + row * n, p, n, lane);
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
        if (!grid_barrier(wk.flags, bar_base + ++barriers, G, &timed_out)) goto fail;
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
        if (!grid_barrier(wk.flags, bar_base + ++barriers, G, &timed_out)) goto fail;
        load_direction_head(zv, wk.z, n);
        double rr, rz_new;
        group_totals2(part_rr, part_rz, G, slots2, &rr, &rz_new);
        if (iterations >= min_iter && sqrt(rr) <= tol * rhs_norm) break;
        if (fabs(rz_old) <= kEps) break;
        const double beta = rz_new / rz_old;
        update_direction(p, zv, wk.z, beta, n, false);
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
// [2n, 2n + nt kTB^2) diagonal-tile inverses, then V = L^-1 (n x n), the
// solve partials (nt x n) and the control words of the factor launch: tile
// flags of L (nt (nt + 1) / 2 ints) and of V (nt (nt - 1) / 2 ints), then 4
// counters.  Only n > kSmallMax uses more than the first 2n.
namespace {
size_t vinv_offset(int n) { return 2 * (size_t)n + tri_blocks(n) * kTB * kTB; }
size_t part_offset(int n) { return vinv_offset(n) + (size_t)n * n; }
size_t control_offset(int n) { return part_offset(n) + tri_blocks(n) * (size_t)n; }
size_t control_doubles(int n) {
  const size_t nt = tri_blocks(n);
  return (nt * nt + 1) / 2 + 2;  // nt (nt + 1) / 2 + nt (nt - 1) / 2 = nt^2 ints, 4 uints
}
}  // namespace

size_t rcs_cholesky_tmp_doubles(int n) {
  if (n <= kSmallMax) return 2 * (size_t)n;
  return control_offset(n) + control_doubles(n);
}

namespace {
FactorWork factor_work(double* tmp, int n) {
  const size_t nt = tri_blocks(n);
  FactorWork fw;
  fw.dinv = tmp + n;
  fw.linv = tmp + 2 * (size_t)n;
  fw.vinv = tmp + vinv_offset(n);
  fw.ready = reinterpret_cast<int*>(tmp + control_offset(n));
  fw.vready = fw.ready + factor_tiles(n);
  fw.ctl = reinterpret_cast<unsigned int*>(fw.ready + nt * nt + (nt & 1));
  return fw;
}

int device_cus(int* cus) {
  int dev = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (*cus < 1) *cus = 1;
  return CT_OK;
}

// x = V^T (V rhs) with the inverse the factor launch left in tmp
int launch_inverse_apply(const double* rhs, double* x, int n, LMState* st, double* tmp, hipStream_t s) {
  const unsigned int tiles = (unsigned int)factor_tiles(n);
  const unsigned int rgrid = (unsigned int)((n + 255) / 256);
  double* y = tmp;
  double* part = tmp + part_offset(n);
  const double* vinv = tmp + vinv_offset(n);
  const double* linv = tmp + 2 * (size_t)n;
  hipLaunchKernelGGL(k_inv_partials<false>, dim3(tiles), dim3(kApplyThreads), 0, s, vinv, linv, rhs, part, n, st);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_inv_reduce<false>, dim3(rgrid), dim3(256), 0, s, part, y, n, st);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_inv_partials<true>, dim3(tiles), dim3(kApplyThreads), 0, s, vinv, linv, y, part, n, st);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_inv_reduce<true>, dim3(rgrid), dim3(256), 0, s, part, x, n, st);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

__global__ void k_zero_words(double* p, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}
}  // namespace

int rcs_cholesky_prepare(double* tmp, int n, hipStream_t s) {
  // zero the flags and counters of the factor launch (a fill kernel, so the
  // call may sit inside a captured graph)
  if (n <= kSmallMax) return CT_OK;
  const size_t words = control_doubles(n);
  const unsigned int grid = (unsigned int)std::min<size_t>((words + 255) / 256, 64);
  hipLaunchKernelGGL(k_zero_words, dim3(grid), dim3(256), 0, s, tmp + control_offset(n), words);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_rcs_cholesky_solve(double* S, const double* rhs, double* x, int n, LMState* st, double* tmp,
                              hipStream_t s, int max_groups) {
  if (n <= 0) return CT_OK;
  if (n <= kSmallMax) {
    hipLaunchKernelGGL(k_chol_small, dim3(1), dim3(kSmallThreads), 0, s, S, rhs, x, tmp, n, st);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }
  int cus = 0;
  CT_TRY(device_cus(&cus));
  const size_t tiles = tri_blocks(n) * tri_blocks(n);  // factor and inverse tiles
  size_t G = std::min<size_t>(tiles, 2 * (size_t)cus);  // two workgroups fit a CU (LDS)
  if (max_groups > 0) {
    // a capped launch shares the chip: two workgroups per tile column keep a
    // single factorisation on its critical path, more only spin (tickets need
    // no co-residency)
    G = std::min<size_t>(G, std::min<size_t>((size_t)max_groups, std::max<size_t>(16, 2 * tri_blocks(n))));
  }
  hipLaunchKernelGGL(k_chol_tiles, dim3((unsigned)G), dim3(kFThreads), 0, s, S, n, factor_work(tmp, n), st);
  CT_LAUNCH_CHECK();
  if (rhs == nullptr) return CT_OK;  // factorisation only
  return launch_inverse_apply(rhs, x, n, st, tmp, s);
}

int launch_rcs_tri_solve(const double* L, const double* rhs, double* x, int n, LMState* st, double* tmp,
                         hipStream_t s) {
  if (n <= 0) return CT_OK;
  if (n > kSmallMax) return launch_inverse_apply(rhs, x, n, st, tmp, s);  // the factor launch left V in tmp
  hipLaunchKernelGGL(k_tri_small, dim3(1), dim3(kSmallThreads), 0, s, L, rhs, x, tmp, n, st);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

// PCG work layout (doubles): z [n], partials [3 kPcgMaxGroups], the epoch
// word (8 doubles reserved), barrier flags [kPcgMaxGroups] (64-bit each).
size_t rcs_pcg_work_doubles(int n) { return (size_t)n + 4 * (size_t)kPcgMaxGroups + 8; }

namespace {
unsigned long long* rcs_pcg_flags(double* work, int n) {
  return reinterpret_cast<unsigned long long*>(work + n + 3 * (size_t)kPcgMaxGroups + 8);
}
}  // namespace

int rcs_pcg_prepare(double* work, int n, hipStream_t s) {
  // zero the epoch word and the barrier flags once per plan (a fill kernel)
  hipLaunchKernelGGL(k_zero_words, dim3(4), dim3(256), 0, s, work + n + 3 * (size_t)kPcgMaxGroups,
                     8 + (size_t)kPcgMaxGroups);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

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
  wk.epoch = rcs_pcg_barrier(work, n);
  wk.flags = rcs_pcg_flags(work, n);
  hipLaunchKernelGGL(k_pcg, dim3(G), dim3(kPcgThreads), 0, s, S, rhs, x, n, wk, st, (unsigned int)G);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct
