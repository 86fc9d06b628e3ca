// Dense damped normal equations on the device: the computation behind the
// reference's linear-solver plugin names sparse_normal_cholesky /
// sparse_cholesky / sparse_schur / block_schur (cuda_backends.py:158-307,
// ceres_torch_cuda.cpp:3-16) for a dense row-major A [m x n]:
//   H = A^T A + diag(damping),  rhs = A^T b,  H x = rhs   (x = H^-1 A^T b)
// H is formed by a tiled Gram kernel and solved with the FP64 Cholesky of the
// reduced-camera-system solver.  Also the standalone factor / solve entry
// points of that Cholesky, and a variant of the dense solve that allocates
// nothing (its workspace is a static device array).
#include <mutex>
#include <vector>

#include "ba_kernels.h"

namespace ct {

namespace {

constexpr int kGT = 64;   // Gram tile
constexpr int kGK = 16;   // rows of A per LDS stage

// Lower 64x64 tiles of H = A^T A + diag(d); the mirror of every off-diagonal
// entry is written too, so H is stored whole.  256 threads x (4x4).
__global__ __launch_bounds__(256) void k_gram(const double* __restrict__ A, int64_t m, int n, int64_t lda,
                                              const double* __restrict__ damping, double* __restrict__ H) {
  __shared__ double Ai[kGK][kGT + 1];
  __shared__ double Aj[kGK][kGT + 1];
  const int tt = blockIdx.x;
  int ti = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > tt) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
  const int tj = tt - ti * (ti + 1) / 2;
  const int i0 = ti * kGT, j0 = tj * kGT;
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int64_t k0 = 0; k0 < m; k0 += kGK) {
    for (int idx = tid; idx < kGK * kGT; idx += 256) {
      const int r = idx / kGT, c = idx % kGT;
      const int64_t k = k0 + r;
      Ai[r][c] = (k < m && i0 + c < n) ? A[k * lda + i0 + c] : 0.0;
      Aj[r][c] = (k < m && j0 + c < n) ? A[k * lda + j0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kGK; ++r) {
      double pa[4], pb[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        pa[a] = Ai[r][tr + 16 * a];
        pb[a] = Aj[r][tc + 16 * a];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += pa[a] * pb[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + tr + 16 * a, j = j0 + tc + 16 * b;
      if (i < n && j <= i) {
        double v = acc[a][b];
        if (i == j && damping) v += damping[i];
        H[(size_t)i * n + j] = v;
        if (i != j) H[(size_t)j * n + i] = v;
      }
    }
  }
}

// rhs = A^T b, one thread per column (rows read coalesced across columns)
__global__ __launch_bounds__(256) void k_atb(const double* __restrict__ A, int64_t m, int n, int64_t lda,
                                             const double* __restrict__ b, double* __restrict__ rhs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t k = 0; k < m; ++k) s += A[k * lda + i] * b[k];
  rhs[i] = s;
}

__global__ void k_clear_state(LMState* st) {
  if (threadIdx.x == 0) {
    st->done = 0;
    st->skip_step = 0;
    st->step_invalid = 0;
    st->linear_iterations = 0;
  }
}

struct DenseLayout {
  double* H = nullptr;
  double* rhs = nullptr;
  double* tmp = nullptr;
  LMState* st = nullptr;
  size_t bytes = 0;
};

DenseLayout dense_layout(int n, void* base) {
  Carver cv(base);
  DenseLayout L;
  L.H = cv.take<double>((size_t)n * n);
  L.rhs = cv.take<double>((size_t)n);
  L.tmp = cv.take<double>(rcs_cholesky_tmp_doubles(n) + 1);
  L.st = cv.take<LMState>(1);
  L.bytes = cv.offset + 256;
  return L;
}

DenseLayout factor_layout(int n, void* base) {
  Carver cv(base);
  DenseLayout L;
  L.tmp = cv.take<double>(rcs_cholesky_tmp_doubles(n) + 1);
  L.st = cv.take<LMState>(1);
  L.bytes = cv.offset + 256;
  return L;
}

// The status word written by the Cholesky kernels, read after the stream drains.
int read_status(LMState* st, hipStream_t s) {
  int invalid = 0;
  CT_HIP(hipMemcpyAsync(&invalid, &st->step_invalid, sizeof(int), hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  if (invalid) {
    set_error("matrix is not positive definite (a Cholesky pivot was not positive and finite)");
    return CT_ERR_NOT_SPD;
  }
  return CT_OK;
}

int check_dense_args(const double* A, int64_t m, int64_t n, int64_t lda, const double* b, const double* x) {
  CT_CHECK_ARG(m >= 0 && n >= 0, "negative size");
  CT_CHECK_ARG(n <= 65536, "n too large for a dense solve");
  CT_CHECK_ARG(lda >= n, "lda must be >= n");
  CT_CHECK_ARG(n == 0 || x != nullptr, "x is null");
  CT_CHECK_ARG(m == 0 || n == 0 || (A != nullptr && b != nullptr), "A / b are null");
  return CT_OK;
}

}  // namespace

int launch_normal_equations(const double* A, int64_t m, int n, int64_t lda, const double* b, const double* damping,
                            double* x, void* workspace, hipStream_t s) {
  DenseLayout L = dense_layout(n, workspace);
  hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(64), 0, s, L.st);
  CT_LAUNCH_CHECK();
  CT_TRY(rcs_cholesky_prepare(L.tmp, n, s));
  const int T = (n + kGT - 1) / kGT;
  hipLaunchKernelGGL(k_gram, dim3(T * (T + 1) / 2), dim3(256), 0, s, A, m, n, lda, damping, L.H);
  CT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_atb, dim3((n + 255) / 256), dim3(256), 0, s, A, m, n, lda, b, L.rhs);
  CT_LAUNCH_CHECK();
  return launch_rcs_cholesky_solve(L.H, L.rhs, x, n, L.st, L.tmp, s);
}

// ---------------------------------------------------------------------------
// No-allocation dense solve: the workspace of ct_normal_equations_solve_f64
// lives in this file-scope device array instead of a caller's buffer.  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): the device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr int64_t kNoAllocMaxN = CT_NOALLOC_MAX_N;
constexpr size_t kNoAllocScratchBytes = (size_t)4608 * 1024;  // >= the dense layout at n = 512 (4,502,624 bytes)
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on the first call on that device).
int scratch_address(void** out) {
  int dev = -1;
  CT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index out of range for the no-allocation scratch table");
    return CT_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return CT_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CT_OK;
}

}  // namespace

}  // namespace ct

using namespace ct;

extern "C" {

int ct_dense_workspace_bytes(int64_t n, size_t* bytes) {
  clear_error();
  CT_CHECK_ARG(bytes != nullptr && n >= 0 && n <= 65536, "bad argument");
  *bytes = dense_layout((int)n, nullptr).bytes;
  return CT_OK;
}

int ct_normal_equations_solve_f64(const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                                  const double* damping, double* x, void* workspace, size_t workspace_bytes,
                                  ct_stream_t stream) {
  clear_error();
  CT_TRY(check_dense_args(A, m, n, lda, b, x));
  if (n == 0) return CT_OK;
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= dense_layout((int)n, nullptr).bytes,
               "workspace too small (ct_dense_workspace_bytes)");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  CT_TRY(launch_normal_equations(A, m, (int)n, lda, b, damping, x, workspace, s));
  return read_status(dense_layout((int)n, workspace).st, s);
}

int ct_normal_equations_solve_f64_noalloc(const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                                          const double* damping, double* x, ct_stream_t stream) {
  clear_error();
  CT_TRY(check_dense_args(A, m, n, lda, b, x));
  if (n > kNoAllocMaxN) {
    set_error("n = " + std::to_string(n) + " exceeds the no-allocation limit of " + std::to_string(kNoAllocMaxN) +
              " unknowns (use ct_normal_equations_solve_f64 with a workspace)");
    return CT_ERR_ARGUMENT;
  }
  if (n == 0) return CT_OK;
  if (dense_layout((int)n, nullptr).bytes > kNoAllocScratchBytes) {
    set_error("no-allocation scratch too small for n = " + std::to_string(n));
    return CT_ERR_ARGUMENT;
  }
  void* ws = nullptr;
  CT_TRY(scratch_address(&ws));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the state word and the Cholesky's control words are cleared by fill
  // kernels inside launch_normal_equations; H and rhs are written whole
  CT_TRY(launch_normal_equations(A, m, (int)n, lda, b, damping, x, ws, s));
  return read_status(dense_layout((int)n, ws).st, s);
}

int ct_block_schur_solve_f64(const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                             int64_t num_eliminate, const double* damping, double* x, void* workspace,
                             size_t workspace_bytes, ct_stream_t stream) {
  // The Schur complement of the first num_eliminate unknowns of an SPD system
  // has the same solution as the whole system; one Cholesky of H gives it.
  clear_error();
  CT_CHECK_ARG(num_eliminate >= 0, "num_eliminate must be >= 0");
  return ct_normal_equations_solve_f64(A, m, n, lda, b, damping, x, workspace, workspace_bytes, stream);
}

int ct_rcs_workspace_bytes(int64_t n, size_t* bytes) {
  clear_error();
  CT_CHECK_ARG(bytes != nullptr && n >= 0 && n <= 65536, "bad argument");
  *bytes = factor_layout((int)n, nullptr).bytes;
  return CT_OK;
}

int ct_rcs_potrf_f64(double* S, int64_t n, void* workspace, size_t workspace_bytes, ct_stream_t stream) {
  clear_error();
  CT_CHECK_ARG(n >= 0 && n <= 65536, "bad size");
  if (n == 0) return CT_OK;
  CT_CHECK_ARG(S != nullptr, "S is null");
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= factor_layout((int)n, nullptr).bytes,
               "workspace too small (ct_rcs_workspace_bytes)");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  DenseLayout L = factor_layout((int)n, workspace);
  hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(64), 0, s, L.st);
  CT_LAUNCH_CHECK();
  CT_TRY(rcs_cholesky_prepare(L.tmp, (int)n, s));
  CT_TRY(launch_rcs_cholesky_solve(S, nullptr, nullptr, (int)n, L.st, L.tmp, s));
  return read_status(L.st, s);
}

int ct_rcs_potrs_f64(const double* L_factor, int64_t n, const double* rhs, double* x, void* workspace,
                     size_t workspace_bytes, ct_stream_t stream) {
  clear_error();
  CT_CHECK_ARG(n >= 0 && n <= 65536, "bad size");
  if (n == 0) return CT_OK;
  CT_CHECK_ARG(L_factor != nullptr && rhs != nullptr && x != nullptr, "null argument");
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= factor_layout((int)n, nullptr).bytes,
               "workspace too small (ct_rcs_workspace_bytes)");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  DenseLayout L = factor_layout((int)n, workspace);
  CT_TRY(launch_rcs_tri_solve(L_factor, rhs, x, (int)n, L.st, L.tmp, s));
  return read_status(L.st, s);
}

}  // extern "C"
