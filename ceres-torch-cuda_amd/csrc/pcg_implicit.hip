// ITERATIVE_SCHUR without the explicit reduced camera system: preconditioned
// CG on S = U + D_c - W^T V^-1 W (iterative_schur_solve, linear.py:702-776,
// preconditioners :462-537, CG :1216-1256) with S p applied from the
// evaluation's per-observation records instead of a stored S.
//
// With the record of observation o holding Jc_o and P_o = Jp_o L_p^-T
// (V_p = L_p L_p^T, damped):
//   t_p     = sum_{o in p} P_o^T (Jc_o p_c(o))                 (point-major)
//   w_o     = Jc_o p_c(o) - P_o t_p(o)
//   (S p)_i = D_i p_i + sum_{o in i} Jc_o^T w_o                 (camera-major)
// which is U_i p_i - sum_o Jc_o^T P_o P_o'^T Jc_o' p, S p written out.
//
// One persistent launch per solve, G co-resident workgroups of 1024 threads,
// two grid barriers per CG iteration:
//   phase 1  wavefronts take the evaluation's point groups (whole points,
//            one lane per observation), form w_o and y_o = Jc_o^T w_o, stage
//            the group's y in LDS and store it in whole lines, in observation
//            order;
//   phase 2  workgroup i sums y over camera i's observations (camera_obs
//            order): (S p)_i and p_i . (S p)_i;
//   then     every workgroup updates its own copy of r and p (in LDS) from
//            the same global values in the same order, so all take the same
//            decisions without a third barrier; workgroup 0 keeps x.
// The block-Jacobi factors of S's diagonal blocks (assembled explicitly by
// the camera-item kernels) are formed once per solve, one block per thread.
#include "ba_kernels.h"
#include "pcg_device.h"

namespace ct {

namespace {

constexpr int kImpThreads = 1024;
constexpr int kImpWaves = kImpThreads / kWave;

__device__ __forceinline__ bool skip_implicit(const LMState* st) { return st->done || st->skip_step; }

// Inverse of each preconditioner block into a.fac (row-major 9x9): of the
// block-Jacobi Cholesky L L^T of S_ii (an LU with partial pivoting when S_ii
// is not SPD), of the diagonal of S_ii (JACOBI, linear.py:531-537) or the
// identity.  Nine threads per camera block: the Cholesky is right-looking,
// thread c updating row c after each pivot; then thread c solves
// L L^T m = e_c for column c of the inverse.  So z = M^-1 r is one 9x9
// product per block in the CG iterations.  kind < 0: a singular block.
// The inverse of the 9x9 block at Sb (row stride n) by LU with partial
// pivoting (the factorisation of lu9, rolled loops: a rare path), worked in
// the LDS arrays W (factors) and M (inverse, row-major).  False if singular.
__device__ __noinline__ bool lu_inverse(const double* Sb, int n, bool split, double* W, double* M, double* dv,
                                        int* piv) {
#pragma unroll 1
  for (int e = 0; e < 81; ++e)
    W[e] = split && ((e / kBlk < 3) != (e % kBlk < 3)) ? 0.0 : Sb[(size_t)(e / kBlk) * n + e % kBlk];
#pragma unroll 1
  for (int j = 0; j < kBlk; ++j) {
    int p = j;
    double best = fabs(W[j * kBlk + j]);
#pragma unroll 1
    for (int r = j + 1; r < kBlk; ++r) {
      const double v = fabs(W[r * kBlk + j]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    piv[j] = p;
    if (!(best > 0.0) || !isfinite(best)) return false;
    if (p != j)
#pragma unroll 1
      for (int k = 0; k < kBlk; ++k) {
        const double t = W[j * kBlk + k];
        W[j * kBlk + k] = W[p * kBlk + k];
        W[p * kBlk + k] = t;
      }
    const double inv = 1.0 / W[j * kBlk + j];
    dv[j] = inv;
#pragma unroll 1
    for (int r = j + 1; r < kBlk; ++r) {
      const double l = W[r * kBlk + j] * inv;
      W[r * kBlk + j] = l;
#pragma unroll 1
      for (int k = j + 1; k < kBlk; ++k) W[r * kBlk + k] -= l * W[j * kBlk + k];
    }
  }
#pragma unroll 1
  for (int cc = 0; cc < kBlk; ++cc) {
    double* zc = M + cc;  // column cc, stride kBlk
#pragma unroll 1
    for (int k = 0; k < kBlk; ++k) zc[k * kBlk] = k == cc ? 1.0 : 0.0;
#pragma unroll 1
    for (int j = 0; j < kBlk; ++j) {
      const int q = piv[j];
      const double t = zc[j * kBlk];
      zc[j * kBlk] = zc[q * kBlk];
      zc[q * kBlk] = t;
    }
#pragma unroll 1
    for (int r = 0; r < kBlk; ++r) {
      double v = zc[r * kBlk];
#pragma unroll 1
      for (int k = 0; k < r; ++k) v -= W[r * kBlk + k] * zc[k * kBlk];
      zc[r * kBlk] = v;
    }
#pragma unroll 1
    for (int r = kBlk - 1; r >= 0; --r) {
      double v = zc[r * kBlk];
#pragma unroll 1
      for (int k = r + 1; k < kBlk; ++k) v -= W[r * kBlk + k] * zc[k * kBlk];
      zc[r * kBlk] = v * dv[r];
    }
  }
  return true;
}

constexpr int kSetupBlocks = 7;  // camera blocks per workgroup (63 of 64 threads)
constexpr int kSetupThreads = 64;

__global__ __launch_bounds__(kSetupThreads) void k_pcg_implicit_setup(ImplicitPcg a) {
  __shared__ double W[kSetupBlocks][81];
  __shared__ double M[kSetupBlocks][81];
  __shared__ double dv[kSetupBlocks][kBlk];
  __shared__ int piv[kSetupBlocks][kBlk];
  __shared__ int ok[kSetupBlocks];
  const LMState* st = a.st;
  if (skip_implicit(st)) return;
  const int n = a.n, precond = st->precond;
  const int lb = threadIdx.x / kBlk, c = threadIdx.x % kBlk;
  const int i = blockIdx.x * kSetupBlocks + lb;
  const bool live = threadIdx.x < kSetupBlocks * kBlk && i < a.C;
  for (int t = blockIdx.x * kSetupThreads + threadIdx.x; t < n; t += gridDim.x * kSetupThreads) a.x[t] = 0.0;
  if (live) {
    // row c of S_ii; with split blocks (quaternion cameras: the rotation and
    // the other six values are separate parameter blocks) the rotation / rest
    // coupling is left out of the block
    const bool split = st->split_blocks != 0;
#pragma unroll
    for (int k = 0; k < kBlk; ++k)
      W[lb][c * kBlk + k] = split && ((c < 3) != (k < 3)) ? 0.0 : a.S[(size_t)(9 * i + c) * n + 9 * i + k];
    if (c == 0) ok[lb] = 1;
  }
  __syncthreads();
  const bool chol = precond != CT_PRECONDITIONER_JACOBI && precond != CT_PRECONDITIONER_IDENTITY;
  if (chol) {
    for (int j = 0; j < kBlk; ++j) {
      if (live && c == j) {
        const double d = W[lb][j * kBlk + j];
        if (!(d > 0.0) || !isfinite(d)) {
          ok[lb] = 0;
        } else {
          const double l = sqrt(d);
          W[lb][j * kBlk + j] = l;
          dv[lb][j] = 1.0 / l;
        }
      }
      __syncthreads();
      if (live && ok[lb] && c > j) W[lb][c * kBlk + j] *= dv[lb][j];
      __syncthreads();
      if (live && ok[lb] && c > j)
        for (int k = j + 1; k <= c; ++k) W[lb][c * kBlk + k] -= W[lb][c * kBlk + j] * W[lb][k * kBlk + j];
      __syncthreads();
    }
  }
  int kind = 3;
  if (live) {
    if (precond == CT_PRECONDITIONER_JACOBI || precond == CT_PRECONDITIONER_IDENTITY) {
      const double dc = precond == CT_PRECONDITIONER_JACOBI ? 1.0 / fmax(W[lb][c * kBlk + c], kEps) : 1.0;
#pragma unroll
      for (int k = 0; k < kBlk; ++k) M[lb][c * kBlk + k] = k == c ? dc : 0.0;
      kind = precond == CT_PRECONDITIONER_JACOBI ? 2 : 3;
    } else if (ok[lb]) {
      // column c of (L L^T)^-1: forward then backward substitution
      double z[kBlk];
#pragma unroll
      for (int r = 0; r < kBlk; ++r) {
        double v = r == c ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < r; ++k) v -= W[lb][r * kBlk + k] * z[k];
        z[r] = v * dv[lb][r];
      }
#pragma unroll
      for (int r = kBlk - 1; r >= 0; --r) {
        double v = z[r];
#pragma unroll
        for (int k = r + 1; k < kBlk; ++k) v -= W[lb][k * kBlk + r] * z[k];
        z[r] = v * dv[lb][r];
      }
#pragma unroll
      for (int r = 0; r < kBlk; ++r) M[lb][r * kBlk + c] = z[r];
      kind = 1;
    } else if (c == 0) {
      // not SPD (rare): the inverse from an LU factorisation with partial pivoting, one thread
      kind = lu_inverse(a.S + (size_t)(9 * i) * n + 9 * i, n, st->split_blocks != 0, W[lb], M[lb], dv[lb], piv[lb])
                 ? 0
                 : -1;
    }
  }
  __syncthreads();
  if (live) {
#pragma unroll
    for (int k = 0; k < kBlk; ++k) a.fac[81 * (size_t)i + c * kBlk + k] = M[lb][c * kBlk + k];
    if (c == 0) a.kind[i] = kind;
  }
}

// z = M^-1 r on the rows this thread owns (i = tid, tid + kImpThreads, ...),
// with its partial sums of r.r and r.z (zero z when some block is singular).
__device__ __forceinline__ void precondition_rows(const ImplicitPcg& a, const double* r, double* z, bool singular,
                                                  double* rr, double* rz) {
  double sr = 0.0, sz = 0.0;
  for (int i = threadIdx.x; i < a.n; i += kImpThreads) {
    const int b = i / kBlk, row = i % kBlk;
    const double* M = a.fac + 81 * (size_t)b + kBlk * row;
    const double* rb = r + kBlk * b;
    double m[kBlk];
#pragma unroll
    for (int k = 0; k < kBlk; ++k) m[k] = M[k];
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < kBlk; ++k) v = fma(m[k], rb[k], v);
    if (singular) v = 0.0;
    z[i] = v;
    sr += r[i] * r[i];
    sz += r[i] * v;
  }
  *rr = sr;
  *rz = sz;
}

__global__ __launch_bounds__(kImpThreads) void k_pcg_implicit(ImplicitPcg a, unsigned int G) {

  extern __shared__ double vec[];  // p, r, z [n] each: identical in every workgroup
  __shared__ double seg[kImpWaves][kWave * 9];  // per lane P^T u (3), then y (9), observation order
  __shared__ double red[kImpThreads / kWave];
  __shared__ double red9[kImpWaves][9];
  __shared__ double tot9[9];
  __shared__ double slots[2];
  __shared__ int timed_out, singular_s;
  LMState* st = a.st;
  if (skip_implicit(st)) return;
  const int n = a.n, C = a.C;
  double* p = vec;
  double* r = vec + n;
  double* z = vec + 2 * (size_t)n;
  const int min_iter = st->min_lin_iter, max_iter = st->max_lin_iter;
  const double tol = st->eta;
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const unsigned int g = blockIdx.x;
  const unsigned long long bar_base =
      (unsigned long long)__hip_atomic_load(a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) << 32;
  unsigned int barriers = 0;
  int iterations = 0;
  if (tid == 0) {
    timed_out = 0;
    singular_s = 0;
  }
  // block-Jacobi factors of the diagonal blocks owned by this workgroup
  {
    // every workgroup keeps r, z = M^-1 r and p
    for (int i = tid; i < n; i += kImpThreads) r[i] = a.rhs[i];
    if (tid < C && a.kind[tid] < 0) singular_s = 1;
    __syncthreads();
    const bool singular = singular_s != 0;
    double rz_old = 0.0, rhs_norm = 0.0;
    // it = -1 forms the first z and p; it >= 0 are the CG iterations
    for (int it = -1; it < max_iter; ++it) {
      if (it >= 0) {
        // phase 1: point groups, one lane per observation
        for (int grp = (int)g * kImpWaves + wave; grp < a.num_groups; grp += (int)G * kImpWaves) {
          const int p0 = a.point_groups[grp], p1 = a.point_groups[grp + 1];
          const int ob = a.point_offsets[p0];
          const int cnt = a.point_offsets[p1] - ob;
          const bool live = lane < cnt;
          const int o = ob + lane;
          double u0 = 0.0, u1 = 0.0;
          int head = 0, k = 0;
          const double* rec = a.JP + (size_t)kJPStride * (live ? o : 0);
          if (live) {
            const int pt = a.obs_point[o];
            const int po = a.point_offsets[pt];
            head = po - ob;
            k = a.point_offsets[pt + 1] - po;
            const double* pc = p + 9 * a.obs_camera[o];
#pragma unroll
            for (int j = 0; j < 9; ++j) {
              u0 = fma(rec[j], pc[j], u0);
              u1 = fma(rec[9 + j], pc[j], u1);
            }
#pragma unroll
            for (int s = 0; s < 3; ++s) seg[wave][3 * lane + s] = rec[18 + s] * u0 + rec[21 + s] * u1;
          }
          lds_wave_sync();
          if (live && lane == head) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            for (int q = head; q < head + k; ++q) {
              t0 += seg[wave][3 * q];
              t1 += seg[wave][3 * q + 1];
              t2 += seg[wave][3 * q + 2];
            }
            seg[wave][3 * head] = t0;
            seg[wave][3 * head + 1] = t1;
            seg[wave][3 * head + 2] = t2;
          }
          lds_wave_sync();
          double w0 = 0.0, w1 = 0.0;
          if (live) {
            // the record again (an L2 hit): holding it across the point sums costs registers
            __asm__ volatile("" ::: "memory");
            const double t0 = seg[wave][3 * head], t1 = seg[wave][3 * head + 1], t2 = seg[wave][3 * head + 2];
            w0 = u0 - (rec[18] * t0 + rec[19] * t1 + rec[20] * t2);
            w1 = u1 - (rec[21] * t0 + rec[22] * t1 + rec[23] * t2);
          }
          lds_wave_sync();  // every lane has read its point's total before y overwrites them
          if (live) {
#pragma unroll
            for (int j = 0; j < 9; ++j) seg[wave][9 * lane + j] = rec[j] * w0 + rec[9 + j] * w1;
          }
          lds_wave_sync();
          // the group's y (cnt x 9 doubles, contiguous in observation order) in whole lines
          {
            double* yg = a.y + 9 * (size_t)ob;
            for (int e = lane; e < 9 * cnt; e += kWave) yg[e] = seg[wave][e];
          }
          lds_wave_sync();
        }
        if (!grid_barrier(a.flags, bar_base + ++barriers, G, &timed_out)) goto fail;
        // phase 2: camera i's slots in order -> (S p)_i and p_i . (S p)_i
        for (int i = (int)g; i < C; i += (int)G) {
          const int s0 = a.camera_offsets[i], s1 = a.camera_offsets[i + 1];
          double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          for (int s = s0 + tid; s < s1; s += kImpThreads) {
            const double* y = a.y + 9 * (size_t)a.camera_obs[s];
#pragma unroll
            for (int j = 0; j < 9; ++j) acc[j] += y[j];
          }
          // wave sums, then the waves' sums in order: one barrier for all nine
#pragma unroll
          for (int j = 0; j < 9; ++j) acc[j] = wave_sum(acc[j]);
          if (lane == 0)
#pragma unroll
            for (int j = 0; j < 9; ++j) red9[wave][j] = acc[j];
          __syncthreads();
          if (tid < 9) {
            double v = 0.0;
            for (int q = 0; q < kImpWaves; ++q) v += red9[q][tid];
            v += a.Dc[9 * i + tid] * p[9 * i + tid];
            a.Ap[9 * i + tid] = v;
            tot9[tid] = p[9 * i + tid] * v;
          }
          __syncthreads();
          if (tid == 0) {
            double pap = 0.0;
            for (int j = 0; j < 9; ++j) pap += tot9[j];
            a.part[i] = pap;
          }
        }
        if (!grid_barrier(a.flags, bar_base + ++barriers, G, &timed_out)) goto fail;
        const double denom = group_total(a.part, (unsigned)C, &slots[0]);
        if (fabs(denom) <= kEps) break;
        const double alpha = rz_old / denom;
        if (g == 0)
          for (int i = tid; i < n; i += kImpThreads) a.x[i] += alpha * p[i];
        for (int i = tid; i < n; i += kImpThreads) r[i] -= alpha * a.Ap[i];
        iterations = it + 1;
        __syncthreads();
      }
      double rr_t = 0.0, rz_t = 0.0;
      precondition_rows(a, r, z, singular, &rr_t, &rz_t);
      double rr = block_sum<kImpThreads>(rr_t, red);
      double rz_new = block_sum<kImpThreads>(rz_t, red);
      if (tid == 0) {
        slots[0] = rr;
        slots[1] = rz_new;
      }
      __syncthreads();
      rr = slots[0];
      rz_new = slots[1];
      __syncthreads();
      double beta = 0.0;
      if (it < 0) {
        rhs_norm = fmax(sqrt(rr), kEps);
        if (min_iter <= 0 && sqrt(rr) <= tol * rhs_norm) break;
      } else {
        if (iterations >= min_iter && sqrt(rr) <= tol * rhs_norm) break;
        if (fabs(rz_old) <= kEps) break;
        beta = rz_new / rz_old;
      }
      // the first direction is z itself: p holds whatever the LDS last held
      // (0 * NaN would poison it), so it is not read then
      if (it < 0) {
        for (int i = tid; i < n; i += kImpThreads) p[i] = z[i];
      } else {
        for (int i = tid; i < n; i += kImpThreads) p[i] = z[i] + beta * p[i];
      }
      __syncthreads();
      rz_old = rz_new;
    }
    if (g == 0 && tid == 0) {
      st->linear_iterations = iterations;
      if (singular) st->step_invalid = 1;
    }
    return;
  }
fail:
  if (tid == 0) st->step_invalid = 1;
}

__global__ void k_implicit_zero_flags(unsigned long long* flags, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) flags[i] = 0ull;
}

}  // namespace

size_t implicit_pcg_doubles(int num_cameras, int num_observations) {
  const size_t C = num_cameras, n = 9 * C;
  // y [N][9], Ap [n], part [C], fac [C][81], kind (as doubles), flags, epoch
  return 9 * (size_t)num_observations + n + C + 81 * C + C + kImpMaxGroups + 8;
}

ImplicitPcg implicit_pcg_layout(double* base, int num_cameras, int num_observations) {
  ImplicitPcg a{};
  const size_t C = num_cameras, n = 9 * C;
  double* q = base;
  a.y = q; q += 9 * (size_t)num_observations;
  a.Ap = q; q += n;
  a.part = q; q += C;
  a.fac = q; q += 81 * C;
  a.kind = reinterpret_cast<int32_t*>(q);
  q += C;
  a.flags = reinterpret_cast<unsigned long long*>(q); q += kImpMaxGroups;
  a.epoch = reinterpret_cast<unsigned int*>(q);
  return a;
}

int implicit_pcg_prepare(const ImplicitPcg& a, hipStream_t s) {
  // zero the barrier flags and the epoch word once per plan (a fill kernel)
  hipLaunchKernelGGL(k_implicit_zero_flags, dim3(1), dim3(256), 0, s, a.flags, kImpMaxGroups + 1);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_pcg_implicit(const ImplicitPcg& a, hipStream_t s) {
  if (a.n <= 0) return CT_OK;
  if (a.C > kImplicitMaxCameras) {
    set_error("the implicit reduced-camera PCG supports at most " + std::to_string(kImplicitMaxCameras) + " cameras");
    return CT_ERR_UNSUPPORTED;
  }
  int dev = 0, cus = 0;
  CT_HIP(hipGetDevice(&dev));
  CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int G = cus < 1 ? 1 : cus;  // one workgroup per CU: all co-resident for the grid barriers
  if (G > kImpMaxGroups) G = kImpMaxGroups;
  const size_t lds = 3 * sizeof(double) * (size_t)a.n;
  if (lds > 48 * 1024)
    CT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pcg_implicit),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_pcg_implicit_setup, dim3((a.C + kSetupBlocks - 1) / kSetupBlocks), dim3(kSetupThreads), 0,
                     s, a);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pcg_implicit, dim3(G), dim3(kImpThreads), lds, s, a, (unsigned int)G);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct
