// Dense damped normal equations on the device: the computation behind the
// reference's linear-solver plugin names sparse_normal_cholesky /
// sparse_cholesky / sparse_schur / block_schur (cuda_backends.py:158-307,
// ceres_torch_cuda.cpp:3-16) for a dense row-major A [m x n]:
//   H = A^T A + diag(damping),  rhs = A^T b,  H x = rhs   (x = H^-1 A^T b)
// H and A^T b are formed from row chunks of A on the matrix cores (partial
// tiles, then a fixed-order reduce) and solved with the FP64 Cholesky of the
// reduced-camera-system solver.  Also the standalone factor / solve entry
// points of that Cholesky, and a variant of the dense solve that allocates
// nothing (its workspace is a static device array).
#include <algorithm>
#include <mutex>
#include <vector>

#include "ba_kernels.h"
#include "chol_small.h"

namespace ct {

namespace {

constexpr int kGT = 64;                 // Gram tile (64 x 64 entries of H)
constexpr int kGS = 32;                 // rows of A per LDS stage
constexpr int kGLd = kGT + 16;          // LDS row stride in doubles (see k_gram_partials)
constexpr int kGramPartialTiles = 1024; // partial tiles the workspace holds (chunks x tiles of H)
constexpr int kGramMaxChunks = 64;
constexpr int64_t kGramMinRows = 1024;  // fewest rows of A a chunk is given
constexpr int kXcds = 8;                // workgroups go to the accelerator dies round robin by index

// H = A^T A is split over row chunks of A: workgroup (chunk, tile) forms the
// 64 x 64 tile (ti, tj), tj <= ti, of the chunk's own product on the matrix
// cores and stores it as a partial; k_gram_reduce adds the partials of each
// entry in chunk order, so H does not depend on scheduling.
struct GramShape {
  int T = 0;          // tile rows of H
  int tiles = 0;      // lower tiles, T (T + 1) / 2
  int chunks = 0;     // row chunks of A
  int64_t rows = 0;   // rows per chunk (a multiple of kGS)
};

int gram_chunk_cap(int n) {
  const int T = (n + kGT - 1) / kGT, tiles = T * (T + 1) / 2;
  return std::max(1, std::min(kGramMaxChunks, kGramPartialTiles / std::max(tiles, 1)));
}

// `slots`: workgroups of k_gram_partials resident on the device at once.
// The chunk count is capped so that the grid fits one resident round when
// the tiles allow it (a second, mostly empty round costs as much as a full
// one).
GramShape gram_shape(int64_t m, int n, int slots) {
  GramShape g;
  g.T = (n + kGT - 1) / kGT;
  g.tiles = g.T * (g.T + 1) / 2;
  int64_t c = std::min<int64_t>(gram_chunk_cap(n), std::max<int64_t>(1, (m + kGramMinRows - 1) / kGramMinRows));
  if (c * g.tiles > slots) c = std::max<int64_t>(1, slots / g.tiles);
  g.chunks = (int)c;
  const int64_t per = (m + c - 1) / c;
  g.rows = std::max<int64_t>(kGS, (per + kGS - 1) / kGS * kGS);
  return g;
}

__device__ __forceinline__ void lower_tile(int t, int& ti, int& tj) {
  ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > t) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  tj = t - ti * (ti + 1) / 2;
}

// One partial tile: part[chunk][tile] = A_c[:, i-block]^T A_c[:, j-block]
// over the chunk's rows, as v_mfma_f64_16x16x4f64 blocks.  Wave w owns the
// 32 x 32 quadrant (w / 2, w % 2); in lane l, block b = 2 bi + bj holds in
// item v the entry (32 (w / 2) + 16 bi + 4 v + l / 16, 32 (w % 2) + 16 bj +
// l % 16).  The A operand of an instruction is the 16 x 4 slice
// (lane l: H row l % 16, stage row l / 16), the B operand the 4 x 16 slice
// (lane l: stage row l / 16, H column l % 16); both are rows of the staged
// A.  The LDS row stride of 80 doubles puts the four 16-lane groups of one
// operand read on alternating halves of the banks.
// Work items (chunk, tile) are dealt to the dies (workgroup index mod kXcds)
// in contiguous chunk-major slices, so the tiles of a chunk mostly share a die
// and read its rows of A through that die's L2.
// On a diagonal tile the upper-right quadrant is not needed: that wave forms
// the chunk's part of A^T b for the tile's 64 columns instead.
__global__ __launch_bounds__(256) void k_gram_partials(const double* __restrict__ A, int64_t m, int n, int64_t lda,
                                                       const double* __restrict__ b, GramShape g,
                                                       double* __restrict__ part, double* __restrict__ part_rhs) {
  __shared__ double Ai[kGS][kGLd];
  __shared__ double Aj[kGS][kGLd];
  __shared__ double bs[kGS];
  // die q % kXcds takes the contiguous slice [die * per, die * per + per) of
  // the chunk-major work list; the grid is kXcds * per, so up to kXcds - 1
  // workgroups find no work
  const int q = blockIdx.x, work = g.chunks * g.tiles, per = (work + kXcds - 1) / kXcds;
  const int item = (q % kXcds) * per + q / kXcds;
  if (item >= work) return;
  const int chunk = item / g.tiles, tile = item % g.tiles;
  int ti, tj;
  lower_tile(tile, ti, tj);
  const bool diag = ti == tj;
  const int i0 = ti * kGT, j0 = tj * kGT;
  const int64_t r_begin = (int64_t)chunk * g.rows;
  const int64_t r_end = std::min<int64_t>(m, r_begin + g.rows);
  const int tid = threadIdx.x, w = tid / 64, lane = tid % 64, li = lane & 15, lk = lane >> 4;
  const bool rhs_wave = diag && w == 1;
  const int sr = tid / 64, sc = tid % 64;  // staging: rows sr + 4 u, column sc
  const bool ci_ok = i0 + sc < n, cj_ok = j0 + sc < n;
  dx4 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = dx4{0.0, 0.0, 0.0, 0.0};
  double racc = 0.0;
  double vi[kGS / 4], vj[kGS / 4], vb = 0.0;
  auto fetch = [&](int64_t r0) {
#pragma unroll
    for (int u = 0; u < kGS / 4; ++u) {
      const int64_t r = r0 + 4 * u + sr;
      const bool ok = r < r_end;
      vi[u] = (ok && ci_ok) ? A[r * lda + i0 + sc] : 0.0;
      vj[u] = (!diag && ok && cj_ok) ? A[r * lda + j0 + sc] : 0.0;
    }
    vb = (diag && tid < kGS && r0 + tid < r_end) ? b[r0 + tid] : 0.0;
  };
  if (r_begin < r_end) fetch(r_begin);
  const int pr = 32 * (w >> 1) + li, qc = 32 * (w & 1) + li;
  for (int64_t r0 = r_begin; r0 < r_end; r0 += kGS) {
    __syncthreads();  // the previous stage's readers are done
#pragma unroll
    for (int u = 0; u < kGS / 4; ++u) {
      Ai[4 * u + sr][sc] = vi[u];
      if (!diag) Aj[4 * u + sr][sc] = vj[u];
    }
    if (diag && tid < kGS) bs[tid] = vb;
    __syncthreads();
    if (r0 + kGS < r_end) fetch(r0 + kGS);  // next stage's loads overlap this stage's products
    const double(*Q)[kGLd] = diag ? Ai : Aj;
    if (!rhs_wave) {
#pragma unroll
      for (int k0 = 0; k0 < kGS; k0 += 4) {
        const int k = k0 + lk;
        const double a0 = Ai[k][pr], a1 = Ai[k][pr + 16];
        const double b0 = Q[k][qc], b1 = Q[k][qc + 16];
        acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[3], 0, 0, 0);
      }
    } else {
#pragma unroll 8
      for (int k = 0; k < kGS; ++k) racc = fma(Ai[k][lane], bs[k], racc);
    }
  }
  double* out = part + ((size_t)chunk * g.tiles + tile) * (kGT * kGT);
  if (!rhs_wave) {
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 32 * (w >> 1) + 16 * (bb >> 1) + 4 * v + (lane >> 4);
        const int col = 32 * (w & 1) + 16 * (bb & 1) + (lane & 15);
        out[row * kGT + col] = acc[bb][v];
      }
  } else {
    part_rhs[(size_t)chunk * g.T * kGT + i0 + lane] = racc;
  }
}

// H (both triangles) = sum of the partial tiles in chunk order + diag(d);
// rhs = sum of the partial A^T b in chunk order.  Blocks [0, 16 tiles) take
// 256 entries of a tile each, the rest 256 entries of rhs.
__global__ __launch_bounds__(256) void k_gram_reduce(const double* __restrict__ part,
                                                     const double* __restrict__ part_rhs, int n, GramShape g,
                                                     const double* __restrict__ damping, double* __restrict__ H,
                                                     double* __restrict__ rhs) {
  constexpr int kPerTile = kGT * kGT / 256;
  const int hb = g.tiles * kPerTile;
  if ((int)blockIdx.x < hb) {
    const int tile = blockIdx.x / kPerTile, e = (blockIdx.x % kPerTile) * 256 + threadIdx.x;
    int ti, tj;
    lower_tile(tile, ti, tj);
    const int i = ti * kGT + e / kGT, j = tj * kGT + e % kGT;
    if (i >= n || j > i) return;
    double v = 0.0;
    for (int c = 0; c < g.chunks; ++c) v += part[((size_t)c * g.tiles + tile) * (kGT * kGT) + e];
    if (i == j && damping) v += damping[i];
    H[(size_t)i * n + j] = v;
    if (i != j) H[(size_t)j * n + i] = v;
  } else {
    const int col = (blockIdx.x - hb) * 256 + threadIdx.x;
    if (col >= n) return;
    double v = 0.0;
    for (int c = 0; c < g.chunks; ++c) v += part_rhs[(size_t)c * g.T * kGT + col];
    rhs[col] = v;
  }
}

// Resident workgroups of k_gram_partials on the current device (cached per device).
int gram_slots(int* slots) {
  static std::mutex mu;
  static int cached[64] = {};
  int dev = 0;
  CT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (dev < 0 || dev >= 64) {
    set_error("device index out of range");
    return CT_ERR_ARGUMENT;
  }
  if (cached[dev] == 0) {
    int cus = 0, per_cu = 0;
    CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gram_partials, 256, 0));
    cached[dev] = std::max(1, cus) * std::max(1, per_cu);
  }
  *slots = cached[dev];
  return CT_OK;
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
  double* part = nullptr;      // Gram partial tiles [chunks][tiles][kGT^2]
  double* part_rhs = nullptr;  // partial A^T b [chunks][T kGT]
  LMState* st = nullptr;
  size_t bytes = 0;
};

DenseLayout dense_layout(int n, void* base) {
  Carver cv(base);
  DenseLayout L;
  const size_t T = (size_t)(n + kGT - 1) / kGT, cap = (size_t)gram_chunk_cap(n);
  L.H = cv.take<double>((size_t)n * n);
  L.rhs = cv.take<double>((size_t)n);
  L.tmp = cv.take<double>(rcs_cholesky_tmp_doubles(n) + 1);
  L.part = cv.take<double>(cap * (T * (T + 1) / 2) * kGT * kGT);
  L.part_rhs = cv.take<double>(cap * T * kGT);
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
  int slots = 0;
  CT_TRY(gram_slots(&slots));
  const GramShape g = gram_shape(m, n, slots);
  const unsigned int ggrid = (unsigned)(((g.chunks * g.tiles + kXcds - 1) / kXcds) * kXcds);
  hipLaunchKernelGGL(k_gram_partials, dim3(ggrid), dim3(256), 0, s, A, m, n, lda, b, g, L.part, L.part_rhs);
  CT_LAUNCH_CHECK();
  const unsigned int rgrid = (unsigned)(g.tiles * (kGT * kGT / 256) + (n + 255) / 256);
  hipLaunchKernelGGL(k_gram_reduce, dim3(rgrid), dim3(256), 0, s, L.part, L.part_rhs, n, g, damping, L.H, L.rhs);
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
constexpr size_t kNoAllocScratchBytes = (size_t)36800 * 1024;  // >= the largest dense layout for n <= 512 (n = 512: 37,647,456 bytes)
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
