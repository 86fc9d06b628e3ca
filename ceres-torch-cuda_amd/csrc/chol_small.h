// Dense SPD factorisation of a small system (n <= 192) by one workgroup of
// 8 waves, shared by the reduced-camera solve kernels and the batched
// bundle-adjustment solver (one workgroup per problem).
#pragma once

#include <utility>

#include "ba_device.h"

namespace ct {

typedef double dx4 __attribute__((ext_vector_type(4)));  // one lane's v_mfma_f64_16x16x4f64 result

// v as held by `lane` (a uniform value: read through scalar registers)
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, lane);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(x): the hardware estimate refined by two Newton steps
// y <- y (3/2 - x y^2 / 2) (each roughly doubles the correct bits; the result
// is within a few ulp).  Not positive or not finite x gives NaN or 0.
__device__ __forceinline__ double pivot_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y *= fma(-h * y, y, 1.5);
  y *= fma(-h * y, y, 1.5);
  return y;
}

constexpr int kPW = 16;  // panel width of the diagonal-tile factorisation

// ---------------------------------------------------------------------------
// small n (n <= kSmallMax): one workgroup of 8 waves
//
// The matrix is cut into 16x16 tiles; lower tile t (row-major order) lives in
// the matrix-core accumulator registers of wave t % 8, slot t / 8, for the
// whole factorisation.  The right-hand side rides along as row n, so the
// factorisation of the augmented matrix leaves y = L^-1 rhs in that row (the
// forward substitution costs nothing extra).  Per 16-column panel p:
//  1. the owners of the panel's tiles write them into the packed lower
//     triangle in LDS (row i starts at i (i + 1) / 2);
//  2. 1-4 waves factor the panel in registers, each holding the 16x16
//     diagonal block in lanes 0-15 (every wave factors it identically) and
//     48 rows below it in lanes 16-63; pivots and columns of the diagonal block
//     reach every lane through scalar registers (v_readlane), so the 16 pivot
//     steps need no barrier; the factored rows go back to LDS;
//  3. every wave subtracts L_ip L_jp^T from its tiles right of the panel with
//     v_mfma_f64_16x16x4f64.
// The packed L then stays in LDS for the backward substitution L^T x = y
// (16-row blocks: one wave solves the block's triangle, then all threads
// update the rows above) and is copied to global memory for later solves.
// ---------------------------------------------------------------------------
constexpr int kSmallMax = 192;
constexpr int kSmallThreads = 512;
constexpr int kSmallWaves = kSmallThreads / kWave;
constexpr int kSmallTileRows = (kSmallMax + 1 + 15) / 16;                                   // with the rhs row
constexpr int kSmallSlots = (kSmallTileRows * (kSmallTileRows + 1) / 2 + kSmallWaves - 1) / kSmallWaves;
constexpr int kSmallPacked = (kSmallMax + 1) * (kSmallMax + 2) / 2;                          // rows 0..n
constexpr int kBelowPerWave = kWave - kPW;  // panel rows below the diagonal block per factoring wave

__device__ __forceinline__ int packed_index(int i, int j) { return i * (i + 1) / 2 + j; }

// One 16-column panel held in registers, factored in 16 pivot steps: lane l
// owns row l of the panel (its first 16 rows are the diagonal block, which
// every factoring wave holds identically).  Pivots and columns reach every
// lane through scalar registers (v_readlane), so a step needs no LDS round
// trip.  A wave issues in order, so each step updates the next column first
// and takes the next pivot's reciprocal square root before the remaining
// column updates: that root's latency overlaps them instead of starting the
// next step.  The first `valid` columns are real; padding columns get the
// reciprocal 0 and change nothing.  On return lane JJ < 16 holds step JJ's
// reciprocal in `pin`, and `bad` is set if a real pivot was not positive and
// finite.  Steps are expanded at compile time, so a[] is only indexed by
// constants; kFull (every column real) drops the per-step column checks.
template <int JJ, bool kFull>
__device__ __forceinline__ void panel_step_ahead(double (&a)[kPW], double& piv, double& inv, double& pin, int& bad,
                                                 int l, int valid) {
  // piv, inv: this step's pivot and reciprocal square root, taken during the previous step
  if ((kFull || JJ < valid) && (!(piv > 0.0) || !isfinite(piv))) bad = 1;
  if (l == JJ) pin = inv;
  a[JJ] = (l == JJ) ? piv * inv : a[JJ] * inv;
  const double f = l > JJ ? a[JJ] : 0.0;
  if constexpr (JJ + 1 < kPW) {
    a[JJ + 1] = fma(-f, lane_value(f, JJ + 1), a[JJ + 1]);
    const double next = lane_value(a[JJ + 1], JJ + 1);
    const double next_inv = (kFull || JJ + 1 < valid) ? pivot_rsqrt(next) : 0.0;
#pragma unroll
    for (int m = JJ + 2; m < kPW; ++m) a[m] = fma(-f, lane_value(f, m), a[m]);
    piv = next;
    inv = next_inv;
  }
}

template <bool kFull, int... J>
__device__ __forceinline__ void panel_steps_ahead(double (&a)[kPW], double& pin, int& bad, int l, int valid,
                                                  std::integer_sequence<int, J...>) {
  double piv = lane_value(a[0], 0);
  double inv = (kFull || valid > 0) ? pivot_rsqrt(piv) : 0.0;
  (panel_step_ahead<J, kFull>(a, piv, inv, pin, bad, l, valid), ...);
}

// Lower tile t (row-major enumeration of the tiles with bj <= bi) -> (bi, bj).
__device__ __forceinline__ void small_tile_rc(int t, int& bi, int& bj) {
  int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > t) --i;
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  bi = i;
  bj = t - i * (i + 1) / 2;
}

// Backward substitution L^T x = y on the packed factor in LDS, in place in
// y[0:n] (every thread calls it).  dinv holds the pivots' reciprocals.  Per
// 16-row block, wave 0 solves the block's triangle with the factor entries
// and pivots preloaded into registers and the solved values broadcast through
// scalar registers; then every thread updates one row above the block.
__device__ void small_backward(const double* Lp, const double* dinv, double* y, int n) {
  const int tid = threadIdx.x;
  const int nb = (n + kPW - 1) / kPW;
  for (int p = nb - 1; p >= 0; --p) {
    const int b0 = kPW * p, kb = min(kPW, n - b0);
    if (tid < kWave) {
      double lc[kPW], dj[kPW];
#pragma unroll
      for (int j = 0; j < kPW; ++j) {
        lc[j] = (j < kb && tid < j) ? Lp[packed_index(b0 + j, b0 + tid)] : 0.0;  // L(b0 + j, b0 + tid)
        dj[j] = j < kb ? dinv[b0 + j] : 0.0;
      }
      double yl = tid < kb ? y[b0 + tid] : 0.0;
#pragma unroll
      for (int j = kPW - 1; j >= 0; --j) {
        if (j < kb) {
          const double xj = lane_value(yl, j) * dj[j];
          yl = tid == j ? xj : fma(-lc[j], xj, yl);
        }
      }
      if (tid < kb) y[b0 + tid] = yl;
    }
    __syncthreads();
    for (int c = tid; c < b0; c += kSmallThreads) {
      double s = 0.0;
      for (int j = 0; j < kb; ++j) s = fma(Lp[packed_index(b0 + j, c)], y[b0 + j], s);
      y[c] -= s;
    }
    __syncthreads();
  }
}

// Forward substitution L y = b in place in y[0:n] (every thread calls it).
__device__ void small_forward(const double* Lp, const double* dinv, double* y, int n) {
  const int tid = threadIdx.x;
  for (int b0 = 0; b0 < n; b0 += kPW) {
    const int kb = min(kPW, n - b0);
    if (tid < kWave) {
      double lr[kPW], dj[kPW];
#pragma unroll
      for (int j = 0; j < kPW; ++j) {
        lr[j] = (j < tid && tid < kb) ? Lp[packed_index(b0 + tid, b0 + j)] : 0.0;  // L(b0 + tid, b0 + j)
        dj[j] = j < kb ? dinv[b0 + j] : 0.0;
      }
      double yl = tid < kb ? y[b0 + tid] : 0.0;
#pragma unroll
      for (int j = 0; j < kPW; ++j) {
        if (j < kb) {
          const double yj = lane_value(yl, j) * dj[j];
          yl = tid == j ? yj : fma(-lr[j], yj, yl);
        }
      }
      if (tid < kb) y[b0 + tid] = yl;
    }
    __syncthreads();
    for (int i = b0 + kb + tid; i < n; i += kSmallThreads) {
      double s = 0.0;
      for (int j = 0; j < kb; ++j) s = fma(Lp[packed_index(i, b0 + j)], y[b0 + j], s);
      y[i] -= s;
    }
    __syncthreads();
  }
}

// Copies between the lower triangle of a row-major n x n matrix in global
// memory and the packed triangle in LDS: wave w takes rows w, w + 8, ...,
// its lanes the row's columns (no index division; each thread's loads of a
// row batch are issued before its LDS stores).
__device__ __forceinline__ void small_store_lower(double* __restrict__ A, const double* Lp, int n) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int r = w; r < n; r += kSmallWaves)
    for (int c = lane; c <= r; c += kWave) A[(size_t)r * n + c] = Lp[packed_index(r, c)];
}

__device__ __forceinline__ void small_load_lower(double* Lp, const double* __restrict__ A, int n) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  constexpr int kRows = 4;  // rows in flight per wave
  for (int r0 = w * kRows; r0 < n; r0 += kSmallWaves * kRows) {
    double v[kRows][(kSmallMax + kWave - 1) / kWave];
#pragma unroll
    for (int q = 0; q < kRows; ++q)
#pragma unroll
      for (int u = 0; u < (kSmallMax + kWave - 1) / kWave; ++u) {
        const int r = r0 + q, c = lane + kWave * u;
        v[q][u] = (r < n && c <= r) ? A[(size_t)r * n + c] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < kRows; ++q)
#pragma unroll
      for (int u = 0; u < (kSmallMax + kWave - 1) / kWave; ++u) {
        const int r = r0 + q, c = lane + kWave * u;
        if (r < n && c <= r) Lp[packed_index(r, c)] = v[q][u];
      }
  }
}

// LDS of the one-workgroup factorisation.
struct SmallCholShared {
  double Lp[kSmallPacked];                  // packed lower triangle, rows 0..n (row n: the rhs)
  double dinv[kSmallMax];                   // pivots' reciprocals
  int tile_rc[kSmallWaves][kSmallSlots][2];  // (bi, bj) of each wave's slots, -1 if unused
  int bad;                                  // a pivot was not positive and finite
};

// Tile slots per wave for systems of at most n unknowns (the rhs row included).
__host__ __device__ constexpr int small_slots(int n) {
  return (((n + 1 + kPW - 1) / kPW) * ((n + 1 + kPW - 1) / kPW + 1) / 2 + kSmallWaves - 1) / kSmallWaves;
}

// Factor the augmented matrix [[A, .], [rhs^T, .]] (lower triangle of the
// row-major n x n A; rhs may be null): on return (after a barrier) sh.Lp holds
// the packed factor with y = L^-1 rhs in row n, sh.dinv the pivots'
// reciprocals and sh.bad the pivot check.  Every thread of the workgroup
// (kSmallThreads) calls it; n <= kMaxN, whose tiles fill kSlots accumulator
// slots per wave.  kFullPanels adds a copy of the panel steps without the
// per-column checks for panels of 16 real columns (faster, more registers).
template <int kMaxN = kSmallMax, bool kFullPanels = true>
__device__ __forceinline__ void small_factor(const double* __restrict__ A, const double* __restrict__ rhs, int n,
                                             SmallCholShared& sh) {
  constexpr int kSlots = small_slots(kMaxN);
  static_assert(kMaxN <= kSmallMax, "small factorisation limit");
  const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
  const int li = lane & 15, lk = lane >> 4;
  const int nt = (n + 1 + kPW - 1) / kPW;          // tile rows, the rhs row included
  const int ntiles = nt * (nt + 1) / 2;
  const int npanels = (n + kPW - 1) / kPW;
  if (tid == 0) sh.bad = 0;
  // load the owned tiles: lower triangle of A, the rhs as row n, zeros elsewhere.
  // The slots' tile coordinates are read back from LDS at every use, so no
  // address derived from them stays live across the panel loop.
  dx4 acc[kSlots];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    const int t = w + kSmallWaves * s;
    int bi = -1, bj = -1;
    acc[s] = dx4{0.0, 0.0, 0.0, 0.0};
    if (t < ntiles) small_tile_rc(t, bi, bj);
    if (lane == 0) {
      sh.tile_rc[w][s][0] = bi;
      sh.tile_rc[w][s][1] = bj;
    }
    if (t < ntiles) {
      const int c = kPW * bj + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = kPW * bi + 4 * v + lk;
        double val = 0.0;
        if (c < n && c <= r) {
          if (r < n) val = A[(size_t)r * n + c];
          else if (r == n && rhs != nullptr) val = rhs[c];
        }
        acc[s][v] = val;
      }
    }
  }
  for (int p = 0; p < npanels; ++p) {
    const int c0 = kPW * p;
    const int valid = min(kPW, n - c0);
    // 1. the panel's tiles into LDS (rows c0..n, columns c0..c0+15 below the diagonal)
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      if (sh.tile_rc[w][s][1] == p) {
        const int c = c0 + li, r0 = kPW * sh.tile_rc[w][s][0];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + 4 * v + lk;
          if (r <= n && c <= r && c < n) sh.Lp[packed_index(r, c)] = acc[s][v];
        }
      }
    }
    __syncthreads();
    // 2. factor the panel: wave k holds the diagonal block and rows c0 + 16 + 48k + (lane - 16)
    const int below = n - (c0 + kPW) + 1;  // rows below the diagonal block, the rhs row included
    const int fwaves = below > 0 ? (below + kBelowPerWave - 1) / kBelowPerWave : 1;
    double a[kPW];
    int row = -1;
    if (w < fwaves) {
      row = lane < kPW ? c0 + lane : c0 + kPW + kBelowPerWave * w + (lane - kPW);
#pragma unroll
      for (int q = 0; q < kPW; ++q) {
        const int c = c0 + q;
        a[q] = (row <= n && c < n && c <= row) ? sh.Lp[packed_index(row, c)] : 0.0;
      }
    }
    __syncthreads();  // every factoring wave has read the raw diagonal block
    if (w < fwaves) {
      double pin = 0.0;
      int pbad = 0;
      if (kFullPanels && valid == kPW)
        panel_steps_ahead<true>(a, pin, pbad, lane, valid, std::make_integer_sequence<int, kPW>{});
      else
        panel_steps_ahead<false>(a, pin, pbad, lane, valid, std::make_integer_sequence<int, kPW>{});
      if (w == 0) {
        if (lane < valid) sh.dinv[c0 + lane] = pin;
        if (pbad && lane == 0) sh.bad = 1;
      }
      if (row <= n && (lane >= kPW || w == 0)) {
#pragma unroll
        for (int q = 0; q < kPW; ++q) {
          const int c = c0 + q;
          if (c < n && c <= row) sh.Lp[packed_index(row, c)] = a[q];
        }
      }
    }
    __syncthreads();
    // 3. trailing update of the tiles right of the panel
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int bj = sh.tile_rc[w][s][1];
      if (bj > p) {
        const int ra = kPW * sh.tile_rc[w][s][0] + li, rb = kPW * bj + li;
#pragma unroll
        for (int k0 = 0; k0 < kPW; k0 += 4) {
          const int c = c0 + k0 + lk;
          const double av = (ra <= n && c < n) ? -sh.Lp[packed_index(ra, c)] : 0.0;
          const double bv = (rb <= n && c < n) ? sh.Lp[packed_index(rb, c)] : 0.0;
          acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[s], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();
  __syncthreads();
}

}  // namespace ct
