// Device helpers shared by the preconditioned-CG kernels (explicit and
// implicit reduced camera system) and the Cholesky kernels: the cross-
// workgroup barrier, fixed-order sums of per-workgroup partials and the 9x9
// camera-block factorisations of the block-Jacobi preconditioner.
#pragma once

#include "ba_kernels.h"

namespace ct {
namespace {

constexpr int kBlk = 9;
constexpr double kEps = 2.220446049250313e-16;

// Every store this wave issued has completed (before a barrier publishes them).
__device__ __forceinline__ void own_stores_done() { __builtin_amdgcn_s_waitcnt(0); }

// LDS written by some lanes of a wave is visible to the wave's other lanes.
__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}


// Grid barrier over G co-resident workgroups: each workgroup stores the
// barrier's number in its own flag slot (release, with L2 write-back for the
// other XCDs) and one wave polls all slots.  Waits are bounded; a timeout
// sets *timed_out and returns false.
__device__ __forceinline__ bool grid_barrier(unsigned long long* flags, unsigned long long target, unsigned int G,
                                             int* timed_out) {
  own_stores_done();
  __syncthreads();
  // the release store writes L2 back (for the other XCDs) before the flag
  if (threadIdx.x == 0) __hip_atomic_store(&flags[blockIdx.x], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < kWave) {
    long spins = 0;
    for (;;) {
      bool ok = true;
      for (unsigned int i0 = threadIdx.x; i0 < G; i0 += 4 * kWave) {
        unsigned long long fv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const unsigned int i = i0 + t * kWave;
          fv[t] = i < G ? __hip_atomic_load(&flags[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : target;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) ok = ok && fv[t] >= target;
      }
      if (__all(ok)) break;
      if (++spins > (1L << 24)) {
        if (threadIdx.x == 0) *timed_out = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return *timed_out == 0;
}

// Sum of part[0:G] in one fixed order (identical in every workgroup); valid
// in every thread after the call.  A lane's loads are issued four at a time
// before their adds.
__device__ __forceinline__ double group_total(const double* part, unsigned int G, double* slot) {
  if (threadIdx.x < kWave) {
    double v = 0.0;
    for (unsigned int i0 = threadIdx.x; i0 < G; i0 += 4 * kWave) {
      double pv[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const unsigned int i = i0 + t * kWave;
        pv[t] = i < G ? part[i] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) v += pv[t];
    }
    v = wave_sum(v);
    if (threadIdx.x == 0) *slot = v;
  }
  __syncthreads();
  const double t = *slot;
  __syncthreads();
  return t;
}

// group_total of two partial arrays from one round of loads (each summed in
// group_total's order).
__device__ __forceinline__ void group_totals2(const double* a, const double* b, unsigned int G, double* slots,
                                              double* ta, double* tb) {
  if (threadIdx.x < kWave) {
    double va = 0.0, vb = 0.0;
    for (unsigned int i0 = threadIdx.x; i0 < G; i0 += 4 * kWave) {
      double pa[4], pb[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const unsigned int i = i0 + t * kWave;
        pa[t] = i < G ? a[i] : 0.0;
        pb[t] = i < G ? b[i] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        va += pa[t];
        vb += pb[t];
      }
    }
    va = wave_sum(va);
    vb = wave_sum(vb);
    if (threadIdx.x == 0) {
      slots[0] = va;
      slots[1] = vb;
    }
  }
  __syncthreads();
  *ta = slots[0];
  *tb = slots[1];
  __syncthreads();
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

// z = M^-1 r for one block: Cholesky solve (kind 1, in registers), LU solve
// (kind 0, rare: worked in place in z, which must not alias r), the inverse
// diagonal (kind 2, JACOBI) or the identity (kind 3).
__device__ __forceinline__ void block_solve(const double* A, const double* dinv, const int* piv, int kind, const double* r,
                            double* z) {
  if (kind == 2 || kind == 3) {
#pragma unroll
    for (int i = 0; i < kBlk; ++i) z[i] = kind == 2 ? dinv[i] * r[i] : r[i];
    return;
  }
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

}  // namespace
}  // namespace ct
